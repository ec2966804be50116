// common.cpp -- logging, errors, environment, counters and resolution of
// the underlying MPI library's symbols.
#include "internal.hpp"

#include <dlfcn.h>

#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace tempi {

int g_log_level = LOG_WARN;

void log_message(int level, const char *fmt, ...) {
  static const char *names[] = {"ERROR", "WARN", "INFO", "DEBUG", "SPEW"};
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  fprintf(stderr, "[tempi_amd %s] %s\n",
          names[level < 0 ? 0 : (level > 4 ? 4 : level)], buf);
}

static thread_local std::string t_last_error;

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
  TEMPI_LOG(LOG_DEBUG, "error %d: %s", code, buf);
  return code;
}

Env &env() {
  static Env e;
  return e;
}

namespace {
std::once_flag g_env_once;
}
void ensure_env() { std::call_once(g_env_once, read_environment); }

// reference: src/internal/env.cpp:23-107.  Only the knobs of the pack path
// are read; the others belong to subsystems this library does not carry.
void read_environment() {
  Env &e = env();
  e.disable = nullptr != std::getenv("TEMPI_DISABLE");
  e.no_pack = nullptr != std::getenv("TEMPI_NO_PACK");
  e.no_type_commit = nullptr != std::getenv("TEMPI_NO_TYPE_COMMIT");
  if (const char *s = std::getenv("TEMPI_SELECT")) {
    if (0 == strcmp(s, "heuristic")) {
      e.select = TEMPI_SELECT_HEURISTIC;
    } else if (0 == strcmp(s, "model")) {
      e.select = TEMPI_SELECT_MODEL;
    } else if (0 == strcmp(s, "autotune")) {
      e.select = TEMPI_SELECT_AUTOTUNE;
    } else {
      e.select = TEMPI_SELECT_DEFAULT;
    }
  }
  if (const char *s = std::getenv("TEMPI_LOG_LEVEL")) {
    g_log_level = atoi(s);
  }
  // cache directory, same search order as the reference
  if (const char *cd = std::getenv("TEMPI_CACHE_DIR")) {
    e.cache_dir = cd;
  } else if (const char *x = std::getenv("XDG_CACHE_HOME")) {
    e.cache_dir = std::string(x) + "/tempi";
  } else if (const char *h = std::getenv("HOME")) {
    e.cache_dir = std::string(h) + "/.tempi";
  } else {
    e.cache_dir = "/var/tmp";
  }
  // send methods for device data (reference: src/internal/env.cpp:73-86)
  e.datatype_method = TEMPI_METHOD_UNKNOWN;
  if (std::getenv("TEMPI_DATATYPE_ONESHOT")) {
    e.datatype_method = TEMPI_METHOD_ONESHOT;
  }
  if (std::getenv("TEMPI_DATATYPE_DEVICE")) {
    e.datatype_method = TEMPI_METHOD_DEVICE;
  }
  if (std::getenv("TEMPI_DATATYPE_STAGED")) {
    e.datatype_method = TEMPI_METHOD_STAGED;
  }
  if (std::getenv("TEMPI_DATATYPE_AUTO")) {
    e.datatype_method = TEMPI_METHOD_UNKNOWN;
  }
  e.contiguous_method = TEMPI_METHOD_UNKNOWN;
  e.contiguous_auto = false;
  if (std::getenv("TEMPI_CONTIGUOUS_STAGED")) {
    e.contiguous_method = TEMPI_METHOD_STAGED;
  }
  if (std::getenv("TEMPI_CONTIGUOUS_AUTO")) {
    e.contiguous_auto = true;
  }
  e.gpu_aware_mpi = nullptr != std::getenv("TEMPI_MPI_GPU_AWARE");
  if (const char *m = std::getenv("TEMPI_METHOD_MODEL")) {
    e.method_model_path = m;
  } else {
    e.method_model_path.clear();
  }
  if (const char *m = std::getenv("TEMPI_MODEL")) {
    e.model_path = m;
  } else {
    e.model_path.clear();
  }
}

Counters &counters() {
  static Counters c;
  return c;
}

// The next definition of `name` after this object (RTLD_NEXT).  When this
// library was loaded with dlopen and libmpi is not behind it in the lookup
// scope, RTLD_NEXT finds nothing; then the object that defines a function
// this library never overrides (MPI_Type_get_envelope) is opened and
// searched instead -- the same recovery the reference makes with
// dlopen("libmpi.so") (src/internal/symbols.cpp:35-49).
static void *resolve(const char *name, void *self) {
  dlerror();
  void *p = dlsym(RTLD_NEXT, name);
  if (p && p != self) {
    return p;
  }
  Dl_info info{};
  if (0 == dladdr(reinterpret_cast<void *>(&MPI_Type_get_envelope), &info) ||
      nullptr == info.dli_fname) {
    return nullptr;
  }
  void *h = dlopen(info.dli_fname, RTLD_LAZY | RTLD_NOLOAD);
  if (!h) {
    h = dlopen(info.dli_fname, RTLD_LAZY);
  }
  if (!h) {
    return nullptr;
  }
  p = dlsym(h, name);
  return (p && p != self) ? p : nullptr;
}

const RealMpi &real_mpi() {
  static RealMpi r;
  static std::once_flag once;
  std::call_once(once, [] {
#define TEMPI_RESOLVE(field, sym)                                              \
  do {                                                                         \
    void *p_ = resolve(#sym, reinterpret_cast<void *>(&sym));                  \
    if (!p_) {                                                                 \
      fprintf(stderr, "[tempi_amd ERROR] cannot resolve %s in the MPI "        \
                      "library\n", #sym);                                      \
      abort();                                                                 \
    }                                                                          \
    r.field = reinterpret_cast<decltype(r.field)>(p_);                         \
  } while (0)
    TEMPI_RESOLVE(Init, MPI_Init);
    TEMPI_RESOLVE(Init_thread, MPI_Init_thread);
    TEMPI_RESOLVE(Finalize, MPI_Finalize);
    TEMPI_RESOLVE(Type_commit, MPI_Type_commit);
    TEMPI_RESOLVE(Type_free, MPI_Type_free);
    TEMPI_RESOLVE(Pack, MPI_Pack);
    TEMPI_RESOLVE(Unpack, MPI_Unpack);
    TEMPI_RESOLVE(Send, MPI_Send);
    TEMPI_RESOLVE(Recv, MPI_Recv);
    TEMPI_RESOLVE(Isend, MPI_Isend);
    TEMPI_RESOLVE(Irecv, MPI_Irecv);
    TEMPI_RESOLVE(Wait, MPI_Wait);
    TEMPI_RESOLVE(Waitall, MPI_Waitall);
    TEMPI_RESOLVE(Test, MPI_Test);
    TEMPI_RESOLVE(Neighbor_alltoallw, MPI_Neighbor_alltoallw);
    TEMPI_RESOLVE(Comm_free, MPI_Comm_free);
#undef TEMPI_RESOLVE
  });
  return r;
}

} // namespace tempi

using namespace tempi;

TEMPI_EXPORT int tempi_version(void) { return 100; }

TEMPI_EXPORT const char *tempi_last_error(void) {
  return t_last_error.c_str();
}

TEMPI_EXPORT int tempi_counters_get(tempi_counters *out) {
  if (!out) {
    return set_error(TEMPI_ERR_ARG, "null output");
  }
  Counters &c = counters();
  out->pack_kernel_launches = c.pack_kernel_launches;
  out->unpack_kernel_launches = c.unpack_kernel_launches;
  out->batch_launches = c.batch_launches;
  out->graph_launches = c.graph_launches;
  out->library_pack_calls = c.library_pack_calls;
  out->library_unpack_calls = c.library_unpack_calls;
  out->commits_canonical = c.commits_canonical;
  out->commits_unsupported = c.commits_unsupported;
  out->model_inferences = c.model_inferences;
  out->autotune_runs = c.autotune_runs;
  out->verified_choices = c.verified_choices;
  out->p2p_device = c.p2p_device;
  out->p2p_oneshot = c.p2p_oneshot;
  out->p2p_staged = c.p2p_staged;
  out->p2p_library = c.p2p_library;
  out->trial_launches = c.trial_launches;
  out->method_model_choices = c.method_model_choices;
  out->host_staged_calls = c.host_staged_calls;
  out->signalled_waits = c.signalled_waits;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_counters_reset(void) {
  Counters &c = counters();
  c.pack_kernel_launches = 0;
  c.unpack_kernel_launches = 0;
  c.batch_launches = 0;
  c.graph_launches = 0;
  c.library_pack_calls = 0;
  c.library_unpack_calls = 0;
  c.commits_canonical = 0;
  c.commits_unsupported = 0;
  c.model_inferences = 0;
  c.autotune_runs = 0;
  c.verified_choices = 0;
  c.p2p_device = 0;
  c.p2p_oneshot = 0;
  c.p2p_staged = 0;
  c.p2p_library = 0;
  c.trial_launches = 0;
  c.method_model_choices = 0;
  c.host_staged_calls = 0;
  c.signalled_waits = 0;
  return TEMPI_SUCCESS;
}

TEMPI_EXPORT int tempi_mpi_handle_size(void) {
  return (int)sizeof(MPI_Datatype);
}

TEMPI_EXPORT int64_t tempi_mpi_constant(const char *name, int *ok) {
  struct Entry {
    const char *n;
    int64_t v;
  };
#define E(x) {#x, (int64_t)(intptr_t)(x)}
  static const Entry table[] = {
      E(MPI_BYTE),          E(MPI_CHAR),          E(MPI_SHORT),
      E(MPI_INT),           E(MPI_LONG),          E(MPI_LONG_LONG),
      E(MPI_FLOAT),         E(MPI_DOUBLE),        E(MPI_UNSIGNED_CHAR),
      E(MPI_UNSIGNED),      E(MPI_INT8_T),        E(MPI_INT16_T),
      E(MPI_INT32_T),       E(MPI_INT64_T),       E(MPI_UINT8_T),
      E(MPI_UINT64_T),      E(MPI_DOUBLE_INT),    E(MPI_FLOAT_INT),
      E(MPI_PACKED),        E(MPI_COMM_WORLD),    E(MPI_COMM_SELF),
      E(MPI_DATATYPE_NULL), E(MPI_ORDER_C),       E(MPI_ORDER_FORTRAN),
      E(MPI_SUCCESS),       E(MPI_ERR_TRUNCATE),  E(MPI_ERRORS_RETURN),
      E(MPI_ERRORS_ARE_FATAL), E(MPI_THREAD_SINGLE), E(MPI_THREAD_MULTIPLE),
      E(MPI_REQUEST_NULL),  E(MPI_ANY_SOURCE),    E(MPI_ANY_TAG),
      E(MPI_PROC_NULL),     E(MPI_ERR_COUNT),     E(MPI_ERR_BUFFER),     E(MPI_INFO_NULL),
      E(MPI_UNWEIGHTED),    E(MPI_COMM_NULL),
      {"sizeof(MPI_Status)", (int64_t)sizeof(MPI_Status)},
      {"offsetof(MPI_Status,MPI_SOURCE)", (int64_t)offsetof(MPI_Status, MPI_SOURCE)},
      {"offsetof(MPI_Status,MPI_TAG)", (int64_t)offsetof(MPI_Status, MPI_TAG)},
      {"offsetof(MPI_Status,MPI_ERROR)", (int64_t)offsetof(MPI_Status, MPI_ERROR)},
  };
#undef E
  for (const Entry &e : table) {
    if (name && 0 == strcmp(name, e.n)) {
      if (ok) {
        *ok = 1;
      }
      return e.v;
    }
  }
  if (ok) {
    *ok = 0;
  }
  return 0;
}
