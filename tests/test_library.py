"""CPU checks of the product library: it loads, exports every entry point that
include/atg.h declares, and the host-side input synthesis is deterministic."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import atg
from atg import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "atg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(atg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_version_and_default_opts():
    assert b"gfx950" in atg.lib().atg_version()
    o = atg.default_opts()
    assert (o.psa_plus, o.chains, o.section_size, o.num_splits, o.split_method) == (1, 16, 0, 0, 0)


@pytest.mark.parametrize("dist", list(atg.DIST))
def test_generators_deterministic_and_valid(dist):
    a = atg.generate_weights(dist, 100_000, seed=42, param=1.1 if dist == "pareto" else 0.0)
    b = atg.generate_weights(dist, 100_000, seed=42, param=1.1 if dist == "pareto" else 0.0)
    assert np.array_equal(a, b)
    assert np.all(np.isfinite(a)) and np.all(a >= 0) and a.sum() > 0


def test_generator_shapes():
    w = atg.generate_weights("pareto", 1_000_000, seed=42, param=1.1)
    assert w.min() >= 1.0
    # Pareto(1.1): P(w > x) = x^-1.1
    assert abs(np.mean(w > 10.0) - 10.0 ** -1.1) < 0.002
    w = atg.generate_weights("heavy_mix", 1_000_000, seed=42, param=0.9)
    heavy = w > 1.0
    assert heavy.sum() == 10_000
    assert abs(w[heavy].sum() / w.sum() - 0.9) < 1e-9
    w = atg.generate_weights("zipf_shuffled", 10_000, seed=1, param=1.2)
    assert np.allclose(np.sort(w)[::-1], np.arange(1, 10_001) ** -1.2)


def test_generate_rejects_bad_distribution():
    w = np.empty(10)
    assert atg.lib().atg_generate_weights(9999, 10, 1, 0.0, w.ctypes.data_as(C.c_void_p)) == _lib.ATG_EINVAL


def test_plots_harness_lists_the_reference_sweeps():
    from atg import plots
    assert {"build_soa", "thread_number", "split_variants", "with_transfer", "light_vs_heavy",
            "sampler_method", "sampling_speed", "distribution", "rejection_performance"} <= set(plots.SWEEPS)
    import subprocess, sys, os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-m", "atg.plots", "--help"], capture_output=True, text=True,
                       cwd=os.path.join(root, "alias-table-gpu_amd"))
    assert r.returncode == 0 and "build_soa" in r.stdout
