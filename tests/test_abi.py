"""C ABI (CPU): libqxgw.so loads, exports exactly what include/qxgw.h declares, the ctypes
struct layouts match the header, and the host-only grid entry points agree with the oracle."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import gw_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qxgw.h")
LIB = os.path.join(ROOT, "quasix_amd", "libqxgw.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"QXGW_API\s+[\w\s\*]+?\b(qx_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("qx_create", "qx_rho_response_dev", "qx_screened_wc_dev", "qx_wmn_imag_dev",
                 "qx_sigma_cd_dev", "qx_g0w0_cd_dev", "qx_g0w0_cd", "qx_g0w0_cd_ov_dev"):
        assert must in names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libqxgw.so not built")
    import quasix_amd._lib as L
    return L.load()


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(qx_\w+)", out))
    assert set(declared()) == exported
    import quasix_amd._lib as L
    assert set(L.EXPORTED) == exported


def test_struct_layouts_match_header(lib, tmp_path):
    """Compile a tiny C program against the header and compare sizeof/offsetof."""
    import quasix_amd._lib as L
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "qxgw.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(qx_newton_config),
         sizeof(qx_g0w0_desc), sizeof(qx_stats), offsetof(qx_g0w0_desc, newton),
         offsetof(qx_g0w0_desc, npair), offsetof(qx_stats, seconds),
         offsetof(qx_stats, pi_variant), offsetof(qx_stats, resolved_instances),
         sizeof(qx_batch_instance), offsetof(qx_batch_instance, targets),
         offsetof(qx_batch_instance, static_terms), offsetof(qx_g0w0_desc, qp_solver));
  return 0;
}
"""
    c = str(tmp_path / "layout.c")
    exe = str(tmp_path / "layout")
    open(c, "w").write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
    vals = list(map(int, subprocess.run([exe], capture_output=True, text=True,
                                        check=True).stdout.split()))
    assert vals == [C.sizeof(L.NewtonConfig), C.sizeof(L.G0W0Desc), C.sizeof(L.Stats),
                    L.G0W0Desc.newton.offset, L.G0W0Desc.npair.offset, L.Stats.seconds.offset,
                    L.Stats.pi_variant.offset, L.Stats.resolved_instances.offset,
                    C.sizeof(L.BatchInstance), L.BatchInstance.targets.offset,
                    L.BatchInstance.static_terms.offset, L.G0W0Desc.qp_solver.offset]


@pytest.mark.parametrize("n", [1, 2, 7, 32, 48, 64, 100])
def test_cabi_grid_matches_oracle(lib, n):
    from quasix_amd import core
    g = core.gauss_legendre_grid(n)
    x, w = O.gauss_legendre_nodes_weights(n)
    assert np.max(np.abs(g["nodes"] - x)) < 2e-15
    assert np.max(np.abs(g["weights"] - w)) < 2e-15
    f, wt = core.get_scaled_legendre_roots(n, 0.5)
    f2, wt2 = O.get_scaled_legendre_roots(n, 0.5)
    # 1/(1-t) amplifies the 1e-16 node rounding near t -> 1 (x7000 at n = 100)
    assert np.max(np.abs(f - f2) / f2) < 1e-11
    assert np.max(np.abs(wt - wt2) / wt2) < 1e-11


def test_cabi_rejects_bad_arguments(lib):
    x = np.empty(2)
    assert lib.qx_gauss_legendre_grid(0, x.ctypes.data_as(C.c_void_p),
                                      x.ctypes.data_as(C.c_void_p)) == 1
    assert lib.qx_scaled_legendre_roots(3, 0.5, None, None) == 1
    h = C.c_void_p()
    assert lib.qx_create(None, 0) == 1
    assert lib.qx_last_error(None) == b"null handle"
