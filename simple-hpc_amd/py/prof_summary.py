"""Per-kernel summary of rocprofv3 runs: time (kernel trace) and HBM traffic (FETCH_SIZE /
WRITE_SIZE counter passes), calibrated on a copy kernel of known size.

Restates the reference's method (bench/gpu_prof_summary.py:19-42 with the passes of
bench/slurm/gpu_profile.sbatch:51-77) for this backend's kernels and rocprofv3 on gfx950:
 - dispatches are grouped by kernel body (the function a k_one / k_many launcher runs), grid and
   workgroup size, so every multigrid level and tile shape has its own row;
 - durations are End - Start of the kernel-trace rows; counters are KiB per dispatch;
 - FETCH_SIZE and WRITE_SIZE are each scaled by what they read for the copy kernel `k_copy`,
   which moves --copy-bytes each way (bench.py's probe copies one field array: nrows x ld
   doubles), because FETCH_SIZE counts 64-byte requests of 128-byte lines on this device;
 - rate = calibrated (fetch + write) bytes / duration, against the 8 TB/s HBM peak; optional
   algorithmic bytes per dispatch (--algo NAME=BYTES) give the algorithmic rate beside it.

Usage: python prof_summary.py --trace KT_kernel_trace.csv [--fetch F_counter_collection.csv
           --write W_counter_collection.csv --copy-bytes N] [--algo 'k_correct=83886080' ...]
           [--requests R_counter_collection.csv] [--run k_mom_sweep ...] [--csv] [--top 25]
Prints Markdown tables (or, with --csv, the reference's columns: kernel, time_pct,
us_per_launch, MB_read, MB_written, GB_per_s).
"""
import argparse
import collections
import csv
import re
import sys

PEAK_GBS = 8000.0


_TYPES = {"d": "double", "f": "float", "i": "int", "b": "bool", "j": "unsigned"}


def _mangled_targs(s, pos):
    """Template arguments I...E of an Itanium-mangled name from s[pos] == 'I': integer and bool
    literals (Li5E, Lb1E) and builtin types; None for anything else."""
    out = []
    pos += 1
    while pos < len(s) and s[pos] != "E":
        m = re.match(r"L([ibj])(n?\d+)E", s[pos:])
        if m:
            v = m.group(2).replace("n", "-")
            out.append(("true" if v == "1" else "false") if m.group(1) == "b" else v)
            pos += m.end()
        elif s[pos] in _TYPES:
            out.append(_TYPES[s[pos]])
            pos += 1
        else:
            return None
    return out


def mangled_body(s):
    """Short name of a mangled kernel symbol: for k_one / k_many the body after the address-of
    template argument (XadL), else the first length-prefixed identifier starting k_, with its
    template arguments. (The k_one / k_many symbols carry a non-type template parameter of
    function-pointer type, which this image's demanglers leave mangled.)"""
    start = s.find("XadL")
    for m in re.finditer(r"(\d+)(k_[A-Za-z0-9_]+)", s[max(start, 0):]):
        digits, at = m.group(1), max(start, 0) + m.start(2)
        for cut in range(len(digits)):  # the length may follow digits of another identifier
            n = int(digits[cut:])
            if n and len(s) >= at + n and re.fullmatch(r"k_\w+", s[at:at + n]):
                name = s[at:at + n]
                if s[at + n:at + n + 1] == "I":
                    ta = _mangled_targs(s, at + n)
                    if ta is not None:
                        name += "<" + ",".join(ta) + ">"
                return name
    return None


def body_name(full):
    """Short name of a kernel: the body function of k_one / k_many (template arguments kept,
    spaces and namespaces dropped), else the kernel function itself."""
    if full.startswith("_Z"):
        b = mangled_body(full)
        if b:
            return b
    s = full.replace("(anonymous namespace)::", "").replace(" ", "")
    m = re.search(r"&(?:\w+::)*(k_\w+)(<[^>]*>)?", s)  # k_one / k_many: the body argument
    if not m:
        m = re.search(r"(?:^|[:\s])(k_\w+)(<[^()]*>)?\(", s)  # a __global__ kernel of its own
    if not m:
        m = re.search(r"(__amd_rocclr_\w+)", s)
        return m.group(1) if m else s[:60]
    return m.group(1) + re.sub(r"\bsg::", "", m.group(2) or "")


def _int(row, *names):
    for n in names:
        if n in row and row[n] not in ("", None):
            return int(float(row[n]))
    return 0


def _grid(row):
    if "Grid_Size" in row:
        return _int(row, "Grid_Size")
    return _int(row, "Grid_Size_X") * max(1, _int(row, "Grid_Size_Y")) * \
        max(1, _int(row, "Grid_Size_Z"))


def _wg(row):
    if "Workgroup_Size" in row:
        return _int(row, "Workgroup_Size")
    return _int(row, "Workgroup_Size_X") * max(1, _int(row, "Workgroup_Size_Y")) * \
        max(1, _int(row, "Workgroup_Size_Z"))


def read_trace(path):
    """{(name, grid, wg): [durations ns]}"""
    out = collections.defaultdict(list)
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            key = (body_name(r["Kernel_Name"]), _grid(r), _wg(r))
            out[key].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return out


def longest_run(path, prefix):
    """Durations (ns) of the longest run of back-to-back dispatches (in start order) whose kernel
    name starts with `prefix`, with nothing else dispatched between them: a probe's repeated
    launches of one kernel, as opposed to its launches inside the solver's iterations."""
    with open(path, newline="") as fh:
        rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Start_Timestamp"]))
    best, cur = [], []
    for r in rows:
        if body_name(r["Kernel_Name"]).startswith(prefix):
            cur.append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
            if len(cur) > len(best):
                best = list(cur)
        else:
            cur = []
    return best


def read_counter(path):
    """{(name, grid, wg): [KiB per dispatch]} of the one counter in the file."""
    out = collections.defaultdict(list)
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            key = (body_name(r["Kernel_Name"]), _grid(r), _wg(r))
            out[key].append(float(r["Counter_Value"]))
    return out


def read_requests(path):
    """{(name, grid, wg): [bytes read per dispatch]} from one counter run of TCC_EA0_RDREQ_32B,
    TCC_EA0_RDREQ_64B and TCC_EA0_RDREQ_128B (the _sum forms): the L2's read requests to memory,
    each counted at its own size. FETCH_SIZE's expression on this device counts a 128-byte
    request as 64 bytes (TCC_BUBBLE, which it takes for the 128-byte ones, reads 0 here), which is
    why a whole-line copy needs the factor of 2; these three need no calibration."""
    per = {}
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            d = r.get("Dispatch_Id") or r.get("Correlation_Id")
            e = per.setdefault(d, {"key": (body_name(r["Kernel_Name"]), _grid(r), _wg(r))})
            e[r["Counter_Name"].replace("_sum", "")] = float(r["Counter_Value"])
    out = collections.defaultdict(list)
    size = {"TCC_EA0_RDREQ_32B": 32, "TCC_EA0_RDREQ_64B": 64, "TCC_EA0_RDREQ_128B": 128}
    for e in per.values():
        out[e["key"]].append(sum(b * e.get(c, 0.0) for c, b in size.items()))
    return out


def summarise(trace, fetch=None, write=None, copy_bytes=None, algo=None, reads=None):
    """Rows sorted by total time: dict(kernel, grid, wg, calls, us, share, fetch_MB, write_MB,
    traffic_MB, GBs, frac, algo_MB, algo_GBs). With `reads` (read_requests), the fetch bytes are
    the request-size counts instead of the calibrated FETCH_SIZE."""
    algo = algo or {}
    busy = sum(sum(v) for v in trace.values()) or 1
    ff = wf = None
    if fetch and write and copy_bytes:
        kib = copy_bytes / 1024.0
        cf = [v for (n, _, _), v in fetch.items() if n == "k_copy"]
        cw = [v for (n, _, _), v in write.items() if n == "k_copy"]
        if cf and cw:
            ff = kib / (sum(map(sum, cf)) / sum(map(len, cf)))
            wf = kib / (sum(map(sum, cw)) / sum(map(len, cw)))
    if reads is not None and write and wf is None:
        wf = 1.0  # WRITE_SIZE counts 32- and 64-byte write requests at their sizes
    rows = []
    for key, d in trace.items():
        name, grid, wg = key
        us = sum(d) / len(d) / 1e3
        r = {"kernel": name, "grid": grid, "wg": wg, "calls": len(d), "us": us,
             "total_ms": sum(d) / 1e6, "share": sum(d) / busy}
        have_f = key in reads if reads is not None else (ff is not None and key in fetch)
        if have_f and wf is not None and key in write:
            if reads is not None:
                fmb = sum(reads[key]) / len(reads[key]) / 1e6
            else:
                fmb = sum(fetch[key]) / len(fetch[key]) * 1024 * ff / 1e6
            wmb = sum(write[key]) / len(write[key]) * 1024 * wf / 1e6
            r.update(fetch_MB=fmb, write_MB=wmb, traffic_MB=fmb + wmb,
                     GBs=(fmb + wmb) * 1e6 / (us * 1e-6) / 1e9)
            r["frac"] = r["GBs"] / PEAK_GBS
        for pat, b in algo.items():
            if name.startswith(pat):
                r["algo_MB"] = b / 1e6
                r["algo_GBs"] = b / (us * 1e-6) / 1e9
        rows.append(r)
    rows.sort(key=lambda r: -r["total_ms"])
    return rows, (None if reads is not None else ff, wf)


def markdown(rows, factors, top=25, out=sys.stdout):
    ff, wf = factors
    if ff is not None:
        print(f"Counter calibration on k_copy: fetch factor {ff:.3f}, write factor {wf:.3f}.\n",
              file=out)
    elif wf is not None:
        print("Fetch bytes from the read-request size counters (no calibration).\n", file=out)
    cols = ["kernel", "grid (threads)", "workgroup", "dispatches", "avg us", "share"]
    has_c = any("traffic_MB" in r for r in rows)
    has_a = any("algo_MB" in r for r in rows)
    if has_c:
        cols += ["fetch MB", "write MB", "traffic MB", "counter GB/s", "of peak"]
    if has_a:
        cols += ["algorithmic MB", "algorithmic GB/s", "of peak (alg.)"]
    print("| " + " | ".join(cols) + " |", file=out)
    print("|---|" + "---:|" * (len(cols) - 1), file=out)
    for r in rows[:top]:
        cells = [r["kernel"], str(r["grid"]), str(r["wg"]), str(r["calls"]), f"{r['us']:.2f}",
                 f"{100 * r['share']:.1f}%"]
        if has_c:
            if "traffic_MB" in r:
                cells += [f"{r['fetch_MB']:.2f}", f"{r['write_MB']:.2f}", f"{r['traffic_MB']:.2f}",
                          f"{r['GBs']:.0f}", f"{100 * r['frac']:.1f}%"]
            else:
                cells += [""] * 5
        if has_a:
            if "algo_MB" in r:
                cells += [f"{r['algo_MB']:.2f}", f"{r['algo_GBs']:.0f}",
                          f"{100 * r['algo_GBs'] / PEAK_GBS:.1f}%"]
            else:
                cells += [""] * 3
        print("| " + " | ".join(cells) + " |", file=out)


def reference_csv(rows, out=sys.stdout):
    """The columns of the reference's summary (bench/gpu_prof_summary.py:33-41)."""
    w = csv.writer(out)
    w.writerow(["kernel", "time_pct", "us_per_launch", "MB_read", "MB_written", "GB_per_s"])
    for r in rows:
        if "traffic_MB" not in r:
            continue
        w.writerow([r["kernel"], f"{100 * r['share']:.1f}", f"{r['us']:.1f}",
                    f"{r['fetch_MB']:.1f}", f"{r['write_MB']:.1f}", f"{r['GBs']:.0f}"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=float, default=None)
    ap.add_argument("--requests", help="counter CSV of TCC_EA0_RDREQ_{32B,64B,128B}_sum: exact "
                    "read bytes in place of the calibrated FETCH_SIZE")
    ap.add_argument("--algo", action="append", default=[],
                    help="NAME_PREFIX=BYTES algorithmic bytes per dispatch")
    ap.add_argument("--csv", action="store_true")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--run", action="append", default=[], metavar="NAME_PREFIX",
                    help="also report the longest back-to-back run of this kernel (a probe)")
    a = ap.parse_args(argv)
    algo = {}
    for kv in a.algo:
        k, v = kv.split("=")
        algo[k] = float(v)
    trace = read_trace(a.trace)
    fetch = read_counter(a.fetch) if a.fetch else None
    write = read_counter(a.write) if a.write else None
    reads = read_requests(a.requests) if a.requests else None
    rows, factors = summarise(trace, fetch, write, a.copy_bytes, algo, reads)
    if a.csv:
        reference_csv(rows)
    else:
        markdown(rows, factors, a.top)
        for pre in a.run:
            d = longest_run(a.trace, pre)
            if d:
                print(f"\nLongest back-to-back run of {pre}: {len(d)} dispatches, mean "
                      f"{sum(d) / len(d) / 1e3:.2f} us, min {min(d) / 1e3:.2f} us, max "
                      f"{max(d) / 1e3:.2f} us")
    return 0


if __name__ == "__main__":
    sys.exit(main())
