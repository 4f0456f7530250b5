"""Output formats and the Ghia validation, on the CPU.

 - The .npy field writer (field_format = npy) round-trips exactly through numpy's own loader
   (no pickled objects), header aligned as the format specifies.
 - The Ghia data as the reference holds them (postprocess/ghia.py:74-88 self-checks): 17 points a
   centreline, wall values, the Re 1000 minimum.
 - Centreline extraction on a field whose answer is known (postprocess/fields.py:37-47) and the
   error measure on the Ghia data itself (zero) and on a constant offset (the offset).
 - The whole validation pinned against the reference's published numbers: the oracle (bit-exact
   with the reference) runs the 80 x 80, Re 1000 cavity (multicolour + red-black, tolerance 1e-5)
   to convergence in the reference's 3718 iterations, and its centrelines give the published
   rms_u 0.01698, max_u 0.03268, rms_v 0.00825 and max_v 0.01386 (docs/RESULTS.md:10, 329).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import ROOT

import pyoracle as po
import simple_gpu as sg
import validate_ghia as vg


def _write_npy(path, a):
    h = sg.host()
    h.sh_write_npy.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int, C.c_int]
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert h.sh_write_npy(str(path).encode(), a.ctypes.data_as(C.POINTER(C.c_double)),
                          a.shape[0], a.shape[1]) == 0


def test_npy_writer_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    for shape in [(1, 1), (37, 45), (1029, 17)]:
        a = rng.standard_normal(shape) * 10.0 ** rng.integers(-300, 300, shape)
        a.flat[0] = -0.0
        f = tmp_path / "a.npy"
        _write_npy(f, a)
        b = np.load(f, allow_pickle=False)
        assert b.dtype == np.float64 and b.shape == shape and b.flags["C_CONTIGUOUS"]
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
        raw = f.read_bytes()
        assert raw[:8] == b"\x93NUMPY\x01\x00"
        hlen = raw[8] | (raw[9] << 8)
        assert (10 + hlen) % 64 == 0 and raw[10 + hlen - 1:10 + hlen] == b"\n"
        assert len(raw) == 10 + hlen + 8 * a.size


def test_ghia_data():
    for re in (400, 1000):
        y, u = vg.ghia_u(re)
        x, v = vg.ghia_v(re)
        assert len(y) == len(u) == len(x) == len(v) == 17
        assert y[0] == 0.0 and u[0] == 0.0 and y[-1] == 1.0 and u[-1] == 1.0
        assert x[0] == 0.0 and v[0] == 0.0 and x[-1] == 1.0 and v[-1] == 0.0
        assert np.all(np.diff(y) > 0) and np.all(np.diff(x) > 0)
    assert vg.ghia_u(1000)[1].min() == -0.38289
    assert vg.suspect(400, "v") == [0.90625] and vg.suspect(1000, "v") == []
    with pytest.raises(ValueError):
        vg.ghia_u(3200)


def test_centrelines_and_errors():
    ux = np.tile(np.arange(4.0), (4, 1))  # value = column index
    c = vg.centrelines(ux, ux.T.copy())
    assert np.allclose(c["u"][1:-1], 1.5) and np.allclose(c["v"][1:-1], 1.5)
    assert c["u"][0] == 0 and c["u"][-1] == 1 and c["v"][0] == 0 and c["v"][-1] == 0
    assert np.allclose(c["y"][1:-1], [0.125, 0.375, 0.625, 0.875])
    ux3 = np.arange(9.0).reshape(3, 3)  # odd count: the middle column, flipped (lid row first)
    assert np.array_equal(vg.centrelines(ux3, ux3)["u"][1:-1], [7.0, 4.0, 1.0])
    gy, gu = vg.ghia_u(1000)
    assert vg.errors(gy, gu, gy, gu) == (0.0, 0.0)
    rms, mx = vg.errors(gy, gu + 0.01, gy, gu)
    assert abs(rms - 0.01) < 1e-15 and abs(mx - 0.01) < 1e-15
    gx, gv = vg.ghia_v(400)
    bumped = gv.copy()
    bumped[np.argmin(abs(gx - 0.90625))] += 1.0
    assert vg.errors(gx, bumped, gx, gv, vg.suspect(400, "v")) == (0.0, 0.0)


def _fmt5(xs):
    return [f"{x:.5f}" for x in xs]


def _centred(p, u, v):
    H = p.H
    ux = (u[H:H + p.NX, H:H + p.NY] + u[H:H + p.NX, H + 1:H + 1 + p.NY]) / 2
    uy = (v[H:H + p.NX, H + 1:H + 1 + p.NY] + v[H + 1:H + 1 + p.NX, H + 1:H + 1 + p.NY]) / 2
    return ux, uy


def test_validation_matches_published_numbers(tmp_path):
    cfg = open(os.path.join(ROOT, "tests", "golden", "ghia", "re1000_80.cfg")).read()
    cfg = cfg.replace("backend = gpu", "backend = serial")
    p, _ = sg.params_from_config(cfg)
    po.set_threads(max(1, min(8, os.cpu_count() or 1)))
    cpu = po.Oracle(p)
    n, log, conv = cpu.run()
    assert conv and n == 3718  # docs/RESULTS.md:79, 329
    ux, uy = _centred(p, cpu.get("u"), cpu.get("v"))
    # through the files the CLI writes (one as .npy) and the run loader
    d = tmp_path / "re1000_80"
    d.mkdir()
    np.savetxt(d / "ux_centre.csv", ux, delimiter=",", fmt="%.18e")
    _write_npy(d / "uy_centre.npy", uy)
    (d / "run_info.json").write_text(json.dumps({"Re": 1000, "N_cells_x": 80, "iterations": n,
                                                 "converged": True}))
    row = vg.evaluate(str(d))
    print(row, flush=True)
    # rms_u, max_u, rms_v, max_v as published (docs/RESULTS.md:10, 329)
    assert _fmt5(row[5:]) == ["0.01698", "0.03268", "0.00825", "0.01386"]
    assert vg.main([str(d), "--check", "--max-rms-u", "0.025", "--max-rms-v", "0.015"]) == 0
    assert vg.main([str(d), "--check", "--max-rms-u", "0.015", "--max-rms-v", "0.015"]) == 1


def test_profile_summariser(tmp_path):
    """prof_summary groups dispatches by kernel body, grid and workgroup, calibrates the counters
    on k_copy and gives time shares, traffic and rates (synthetic rocprofv3 CSVs)."""
    import prof_summary as ps
    sweep = ("void sg::k_one<sg::(anonymous namespace)::SweepArgs, &sg::(anonymous namespace)::"
             "k_mom_sweep<5, 40, 56, false, 512, true>, 512>(sg::(anonymous namespace)::"
             "SweepArgs, int const*)")
    copy = ("void sg::k_one<sg::(anonymous namespace)::CopyArgs, &sg::(anonymous namespace)::"
            "k_copy, 256>(sg::(anonymous namespace)::CopyArgs, int const*)")
    own = "void sg::(anonymous namespace)::k_expand<5>(sg::(anonymous namespace)::ExpArgs)"
    assert ps.body_name(sweep) == "k_mom_sweep<5,40,56,false,512,true>"
    assert ps.body_name(copy) == "k_copy" and ps.body_name(own) == "k_expand<5>"
    assert ps.body_name("__amd_rocclr_fillBufferAligned.kd") == "__amd_rocclr_fillBufferAligned"
    # the symbols as rocprofv3 reports them (left mangled: a function-pointer template parameter)
    assert ps.body_name("_ZN2sg5k_oneINS_12_GLOBAL__N_19SweepArgsETnPFvRKT_RKNS_3BlkEEXadL_ZNS1_"
                        "11k_mom_sweepILi5ELi40ELi56ELb0ELi512ELb1EEEvRKS2_S8_EELi512EEEvS3_PKi"
                        ) == "k_mom_sweep<5,40,56,false,512,true>"
    assert ps.body_name("_ZN2sg5k_oneINS_12_GLOBAL__N_12DUIdEETnPFvRKT_RKNS_3BlkEEXadL_ZNS1_11k_"
                        "mg_down_sIdLi32ELi24EEEvRKNS2_IS4_EES9_EELi512EEEvS4_PKi"
                        ) == "k_mg_down_s<double,32,24>"
    assert ps.body_name("_ZN2sg5k_oneINS_8StopArgsETnPFvRKT_RKNS_3BlkEEXadL_ZNS_12_GLOBAL__N_111k"
                        "_stop_ruleES1_S4_EELi64EEEvS1_PKi") == "k_stop_rule"
    kt = tmp_path / "kt.csv"
    kt.write_text("Kind,Kernel_Name,Start_Timestamp,End_Timestamp,Workgroup_Size_X,"
                  "Workgroup_Size_Y,Workgroup_Size_Z,Grid_Size_X,Grid_Size_Y,Grid_Size_Z\n"
                  f'KERNEL,"{sweep}",1000,19000,512,1,1,252928,1,1\n'
                  f'KERNEL,"{sweep}",20000,38000,512,1,1,252928,1,1\n'
                  f'KERNEL,"{copy}",40000,44000,256,1,1,262144,1,1\n')
    cols = "Kernel_Name,Grid_Size,Workgroup_Size,Counter_Name,Counter_Value\n"
    fe, wr = tmp_path / "f.csv", tmp_path / "w.csv"
    fe.write_text(cols + f'"{sweep}",252928,512,FETCH_SIZE,20000\n'
                         f'"{sweep}",252928,512,FETCH_SIZE,22000\n'
                         f'"{copy}",262144,256,FETCH_SIZE,4000\n')
    wr.write_text(cols + f'"{sweep}",252928,512,WRITE_SIZE,8000\n'
                         f'"{sweep}",252928,512,WRITE_SIZE,8000\n'
                         f'"{copy}",262144,256,WRITE_SIZE,8000\n')
    rows, (ff, wf) = ps.summarise(ps.read_trace(kt), ps.read_counter(fe), ps.read_counter(wr),
                                  copy_bytes=8000 * 1024.0, algo={"k_mom_sweep": 41.9e6})
    assert ff == 2.0 and wf == 1.0
    r = rows[0]
    assert r["kernel"].startswith("k_mom_sweep") and r["calls"] == 2 and r["us"] == 18.0
    assert abs(r["share"] - 36.0 / 40.0) < 1e-12
    assert abs(r["fetch_MB"] - 21000 * 1024 * 2 / 1e6) < 1e-9
    assert abs(r["GBs"] - (21000 * 2 + 8000) * 1024 / 18e-6 / 1e9) < 1e-6
    assert abs(r["algo_GBs"] - 41.9e6 / 18e-6 / 1e9) < 1e-9
    # exact read bytes from the request-size counters: 32 x n32 + 64 x n64 + 128 x n128
    rq = tmp_path / "r.csv"
    rq.write_text("Dispatch_Id,Kernel_Name,Grid_Size,Workgroup_Size,Counter_Name,Counter_Value\n"
                  f'1,"{sweep}",252928,512,TCC_EA0_RDREQ_32B_sum,10\n'
                  f'1,"{sweep}",252928,512,TCC_EA0_RDREQ_64B_sum,100\n'
                  f'1,"{sweep}",252928,512,TCC_EA0_RDREQ_128B_sum,1000\n'
                  f'2,"{sweep}",252928,512,TCC_EA0_RDREQ_128B_sum,3000\n')
    reads = ps.read_requests(rq)
    assert sorted(reads[("k_mom_sweep<5,40,56,false,512,true>", 252928, 512)]) == \
        [320 + 6400 + 128000, 384000]
    rows2, (ff2, wf2) = ps.summarise(ps.read_trace(kt), None, ps.read_counter(wr), reads=reads)
    assert ff2 is None and wf2 == 1.0
    assert abs(rows2[0]["fetch_MB"] - (134720 + 384000) / 2 / 1e6) < 1e-12
    assert abs(rows2[0]["write_MB"] - 8000 * 1024 / 1e6) < 1e-12
    import io
    buf = io.StringIO()
    ps.reference_csv(rows, buf)
    assert buf.getvalue().splitlines()[0] == "kernel,time_pct,us_per_launch,MB_read,MB_written,GB_per_s"


def test_bench_runner_rows(tmp_path):
    """The bench runner's config and CSV rows (bench/run_bench.py:33-55, 150-185), with a stand-in
    for the CLI run."""
    import run_bench as rb
    text = rb.make_config(96, 40, tmp_path / "o", {"mg_precision": "mixed", "Re": "400"})
    kv = dict(ln.split(" = ") for ln in text.strip().splitlines())
    assert kv["N_cells_x"] == kv["N_cells_y"] == "96" and kv["maxit"] == "40"
    assert kv["error_tolerance"] == "0" and kv["convection_scheme"] == "QUICK"
    assert kv["diffusion_order"] == "4" and kv["Re"] == "400"  # --set wins over the overrides
    assert kv["backend"] == "gpu" and kv["pressure_solver"] == "mg"
    assert kv["relaxation_factor_p"] == "0.2" and kv["mg_precision"] == "mixed"
    assert kv["output_dir"] == str(tmp_path / "o")
    calls = []

    def fake(exe, grid, threads, iters, extra):
        calls.append((grid, threads))
        phases = [("assemble_u", 0.5 * grid), ("solve_p", 1.0 * grid), ("total", 2.0 * grid),
                  ("unaccounted", 0.1)]
        return phases, {"seconds": 1.5 * grid, "iterations": iters, "device": "gfx950 test",
                        "compiler": "c", "git_hash": "h", "converged": False}

    out = tmp_path / "r.csv"
    rows = rb.run([64, 128], 10, 2, [1, 4], "simple", "t", {}, out, runner=fake)
    # repeat-major, the thread list rotated each repeat
    assert calls == [(64, 1), (64, 4), (128, 1), (128, 4), (64, 4), (64, 1), (128, 4), (128, 1)]
    with open(out) as fh:
        head = fh.readline().strip().split(",")
    assert head == rb.FIXED + ["assemble_u", "solve_p"] + rb.TRAILER
    r = rows[2]
    assert r["grid"] == 128 and r["total_seconds"] == f"{192.0:.9f}"
    assert float(r["seconds_per_iter"]) == pytest.approx(19.2)
    assert float(r["cell_updates_per_s"]) == pytest.approx(128 * 128 * 10 / 192.0, rel=1e-6)
    assert r["cpu"] == "gfx950 test" and r["solve_p"] == f"{128.0:.9f}"


def test_bench_runner_write_config(tmp_path):
    """--write-config writes each grid's run config (the run's output_dir beside it) and runs
    nothing, so the same run can be started under a profiler."""
    import run_bench as rb
    d = tmp_path / "cfgs"
    assert rb.main(["--grids", "64", "2048", "--iters", "100", "--write-config", str(d),
                    "--set", "pressure_solver=gs_redblack", "--exe", "/nonexistent"]) == 0
    assert sorted(p.name for p in d.iterdir()) == ["g2048.cfg", "g64.cfg"]
    kv = dict(ln.split(" = ") for ln in (d / "g2048.cfg").read_text().strip().splitlines())
    assert kv["N_cells_x"] == "2048" and kv["maxit"] == "100"
    assert kv["pressure_solver"] == "gs_redblack" and kv["output_dir"] == str(d / "g2048")
