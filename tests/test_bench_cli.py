"""The bench line's contract on the CPU: a single-GPU refusal of --gpus N > 1, the roofline figures
and the kernel names the counter passes select (no GPU needed)."""
import importlib.util
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_the_sharding_plan():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""


def test_roofline_figures():
    b = _bench()
    r = b.roofline(320e6, 0.25, 20, 1024, "mixed")  # 320 MB in 250 us
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["achieved"] - 1280.0) < 0.05
    assert abs(r["frac"] - 0.16) < 1e-4
    assert r["avg_us"] == 250.0 and r["bytes_per_apply"] == 320000000 and r["applies_timed"] == 20
    assert "k_local_solve<double, float, 1024>" in r["kernel"] and "k_dense_top<double, float, NB>" in r["kernel"]


def test_counter_passes_select_the_kernels_of_each_precision():
    b = _bench()
    # rocprofv3's names of the launches of one apply's local solve, per precision
    names = {
        "mixed": ["void s2l::(anonymous namespace)::k_local_solve<double, float, 1024>(s2l::LocalArgs<double, float>)",
                  "void s2l::(anonymous namespace)::k_dense_top<double, float, 3>(s2l::LocalArgs<double, float>, int4 const*)"],
        "single": ["void s2l::(anonymous namespace)::k_local_solve<float, float, 512>(s2l::LocalArgs<float, float>)",
                   "void s2l::(anonymous namespace)::k_dense_top<float, float, 1>(s2l::LocalArgs<float, float>, int4 const*)"],
        "double": ["void s2l::(anonymous namespace)::k_local_solve<double, double, 1024>(s2l::LocalArgs<double, double>)",
                   "void s2l::(anonymous namespace)::k_dense_top<double, double, 2>(s2l::LocalArgs<double, double>, int4 const*)"],
    }
    for prec, kernels in names.items():
        sel = b.local_kernels(prec)
        for k in kernels:
            assert any(s in k for s in sel), (prec, k)
        # no other precision's local-solve kernels are counted
        for other, ks in names.items():
            if other != prec:
                assert not any(any(s in k for s in sel) for k in ks), (prec, other)
        assert not any(s in "k_copy(long, double const*, double*)" for s in sel)
