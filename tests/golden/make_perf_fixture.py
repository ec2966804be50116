"""Write tests/golden/perf_reference_schema.json: a system-performance file in
the layout the reference's writer produces (src/internal/measure_system.cpp:
36-47 through nlohmann::json::dump(2): keys sorted, each measurement an
object {"iid": bool, "time": seconds}).

The numbers are synthetic (latency + bytes / bandwidth, with a penalty for
short blocks) and exist to exercise the reader; they describe no machine.

    python tests/golden/make_perf_fixture.py
"""
import json
import os

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "perf_reference_schema.json")


def curve(lat, bw, iid_every=0):
    return [{"iid": bool(iid_every and i % iid_every == 0), "time": lat + (1 << i) / bw}
            for i in range(24)]


def table(lat, bw, link=None):
    rows = []
    for i in range(9):
        nbytes = 1 << (2 * i + 6)
        row = []
        for j in range(9):
            bl = min(nbytes, 1 << j)
            eff = bw * min(1.0, bl / 64.0)
            if link:
                eff = min(eff, link)
            row.append({"iid": False, "time": lat + nbytes / eff})
        rows.append(row)
    return rows


def main():
    perf = {
        "cudaKernelLaunch": 2.5e-06,
        "d2h": curve(9e-06, 50e9, 3),
        "h2d": curve(9e-06, 50e9),
        "intraNodeCpuCpuPingpong": curve(1.5e-06, 20e9),
        "intraNodeGpuGpuPingpong": curve(8e-06, 100e9),
        "interNodeCpuCpuPingpong": [],
        "interNodeGpuGpuPingpong": [],
        "packDevice": table(1.2e-05, 4e12),
        "unpackDevice": table(1.2e-05, 4e12),
        "packHost": table(1.0e-05, 4e12, link=50e9),
        "unpackHost": table(1.0e-05, 4e12, link=50e9),
    }
    with open(OUT, "w") as fh:
        json.dump(perf, fh, indent=2, sort_keys=True)
        fh.write("\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
