"""The measurement tooling of bench.py on the CPU: the one-GPU refusal, kernel
names from rocprofv3 records, the traffic table built from FETCH_SIZE /
WRITE_SIZE records (scaled by the copy probe, merged over runs with and
without LDS windows, looked up for the build the timed steps ran) and the
timed-steps line of the kernel table, which leaves out iterations launched
past convergence."""
import csv
import json
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_kernel_names():
    assert bench.kernel_name("void sphgpu::(anonymous namespace)::k_df_pass_divergence<true>(sphgpu::Dev, int)") \
        == "k_df_pass_divergence<true>"
    assert bench.kernel_name("sphgpu::(anonymous namespace)::k_scene_reduce(sphgpu::Dev, int, float, int, int)") \
        == "k_scene_reduce"
    assert bench.kernel_name("_ZN6sphgpu12_GLOBAL__N_112k_copy_probeEPK15HIP_vector_typeIfLj4EEPS2_m") \
        == "k_copy_probe"


def _pmc_csv(path, counter, rows):
    """rows: (kernel name, value, duration in ns)"""
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp",
                    "Grid_Size"])
        t = 1000
        for name, value, dur in rows:
            w.writerow([name, counter, value, t, t + dur, 256])
            t += dur + 100


def test_traffic_table_scaled_merged_and_matched(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "traffic_file", lambda workload: tmp_path / f"{workload}_traffic.txt")
    kib = bench.COPY_PROBE_BYTES / 1024.0
    probe = "sphgpu::(anonymous namespace)::k_copy_probe(HIP_vector_type<float, 4u> const*, " \
            "HIP_vector_type<float, 4u>*, unsigned long)"
    win = "void sphgpu::(anonymous namespace)::k_df_pass_divergence<true>(sphgpu::Dev, int)"
    gat = "void sphgpu::(anonymous namespace)::k_df_pass_divergence<false>(sphgpu::Dev, int)"
    # the copy probe reads half its bytes in FETCH_SIZE units: fetch factor 2
    _pmc_csv(tmp_path / "f_on.csv", "FETCH_SIZE",
             [(probe, kib / 2, 50000), (win, 1000.0, 300000), (win, 1200.0, 300000), (win, 5.0, 3000)])
    _pmc_csv(tmp_path / "w_on.csv", "WRITE_SIZE", [(probe, kib, 50000), (win, 100.0, 300000), (win, 100.0, 300000),
                                                   (win, 0.0, 3000)])
    _pmc_csv(tmp_path / "f_off.csv", "FETCH_SIZE", [(probe, kib / 2, 50000), (gat, 400.0, 320000)])
    _pmc_csv(tmp_path / "w_off.csv", "WRITE_SIZE", [(probe, kib, 50000), (gat, 100.0, 320000)])
    assert bench.summarize_pmc(f"{tmp_path}/f_on.csv,{tmp_path}/f_off.csv",
                               f"{tmp_path}/w_on.csv,{tmp_path}/w_off.csv", "wl", "cmd") == 0
    on = bench.read_traffic("wl", "df_pass_divergence", True)
    off = bench.read_traffic("wl", "df_pass_divergence", False)
    # medians over the launches of 10 us or more; fetch scaled by 2, write by 1
    assert on["fetch_kib"] == 1200.0 and on["dispatches"] == 2
    assert on["traffic_bytes"] == round((1200.0 * 2 + 100.0) * 1024)
    assert off["traffic_bytes"] == round((400.0 * 2 + 100.0) * 1024)
    assert bench.read_traffic("wl", "df_pass_velocity", True) is None
    assert bench.read_traffic("other", "df_pass_divergence", True) is None


def test_timed_steps_line_skips_launches_past_convergence(tmp_path, monkeypatch):
    (tmp_path / "profiles").mkdir()
    monkeypatch.setattr(bench, "ROOT", tmp_path)
    name = "void sphgpu::(anonymous namespace)::k_df_pass_divergence<true>(sphgpu::Dev, int)"
    other = "sphgpu::(anonymous namespace)::k_df_integrate(sphgpu::Dev)"
    # working launches of 300 us, one launch past convergence of 22 us after
    # every tenth, and a kernel of another name
    durs = []
    for k in range(60):
        durs.append((name, 300000 + 1000 * (k % 3)))
        if k % 10 == 9:
            durs.append((name, 22000))
        durs.append((other, 40000))
    with open(tmp_path / "trace.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 0
        for n, d in durs:
            w.writerow([n, t, t + d])
            t += d + 500
    assert bench.summarize_trace(None, str(tmp_path / "trace.csv"), "wl", "cmd",
                                 "k_df_pass_divergence<true>:30") == 0
    text = (tmp_path / "profiles" / "wl_kernel_stats.txt").read_text()
    last = text.strip().splitlines()[-1]
    assert last.startswith("# timed steps: last 30 dispatches of k_df_pass_divergence<true>")
    want = sum(300 + (k % 3) for k in range(30, 60)) / 30
    assert f"mean {want:.2f} us" in last
    rows = [ln.split() for ln in text.splitlines() if not ln.startswith("#")]
    assert {r[0] for r in rows} == {"k_df_pass_divergence<true>", "k_df_integrate"}
