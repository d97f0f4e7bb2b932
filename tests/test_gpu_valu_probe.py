"""The roofline's packed-VALU peak bounds what the GPU issues.

The sDTW roofline prices cells at N_CU x 64 lanes x 2 halves x f_clk element
ops, i.e. one packed wave64 instruction per SIMD per 4 cycles (1.0 per
CU-cycle).  `dtwax_amd/bin/valu_probe` times long runs of independent VALU
instructions at the same nominal clock; none of its packed binary16 or DPP
rates may exceed that peak (else the roofline would understate what the
hardware can do), and binary32 adds, which gfx950 issues from two waves side
by side, must come out faster than packed adds (the basis of the binary32
form's ceiling).
"""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "dtwax_amd", "bin", "valu_probe")


def test_valu_probe_within_the_roofline_peak():
    if not os.path.exists(PROBE):
        pytest.fail("dtwax_amd/bin/valu_probe is not built (make -C dtwax_amd/csrc)")
    print("valu_probe 2000 iterations", flush=True)
    out = subprocess.run([PROBE, "2000"], capture_output=True, text=True, timeout=120, check=True).stdout
    rep = json.loads([line for line in out.splitlines() if line.startswith("{")][-1])
    assert rep["n_cu"] > 0 and rep["clock_mhz"] > 0
    rates = rep["rates"]
    for name in ("v_pk_add_f16", "v_pk_min_f16", "v_pk_mul_f16", "v_pk_fma_f16", "v_mov_b32_dpp"):
        assert len(rates[name]) == 4
        assert all(0.3 < r <= 1.05 for r in rates[name]), (name, rates[name])
    # two waves per SIMD: binary32 adds issue side by side, packed ops do not
    assert rates["v_add_f32"][1] > 1.2 * rates["v_pk_add_f16"][1], rates
