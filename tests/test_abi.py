"""The C-ABI library loads on the CPU and exports every entry point declared
in include/nnmerge.h (no device calls are made here)."""
import ctypes
import os
import re

import nnm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "nnmerge.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nnm_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "nnm_create" in syms and "nnm_nndescent" in syms and "nnm_bruteforce_knn" in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nnm.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(nnm.exported_symbols()) == declared_symbols()


def test_default_params_are_reference_constants():
    p = nnm.default_params()
    assert p.K == 20 and p.rev_cap == 100 and abs(p.delta - 0.002) < 1e-9 and p.rho == 1.0
    assert p.max_iters == 0 and p.rule == nnm.RULE_ALL and p.prec == nnm.PREC_FP32
    assert abs(p.learn_tau - 0.6) < 1e-6 and p.learn_rho == 0.0


def test_params_struct_layout_matches_header():
    # nnm_params: 14 x 4-byte fields then a uint64 seed (offset 56), size 64
    assert ctypes.sizeof(nnm.Params) == 64
    assert nnm.Params.seed.offset == 56
    # ... redo_rows, learned_dists, redo_rounds + pad, deliver_secs
    assert ctypes.sizeof(nnm.Stats) == 8 + 6 * 8 + 3 * 8 + 8 + 16 + 32 + 16 + 16 + 8 + 8 + 8 + 8
    assert nnm.Stats.learned_dists.offset == 176
    assert nnm.Stats.deliver_secs.offset == 192
