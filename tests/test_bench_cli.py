"""Host-side checks of bench.py that need no GPU: the one-GPU contract and
the workload descriptions it reports."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""


def test_workload_descriptions_match_the_workloads():
    import bench

    class Args:
        replicas = 2
        mols_per_replica = 50
        steps = 1

    for name in ("free_diffusion_replicas", "bimol_box", "free_diffusion_box"):
        Args.workload = name
        wl, cfg = bench.build_workload(Args)
        assert cfg["molecules"] == wl.num_molecules
        assert cfg["workload"]
        if "walls" in cfg:
            assert cfg["walls"] == len(wl.walls)
