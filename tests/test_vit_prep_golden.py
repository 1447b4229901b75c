"""The product's host restatement of vf_conversion (msvh_vit_convert in
cudahmmsearch_amd/csrc/msv_host.cpp, p7_oprofile.c:653-740 over
p7_ProfileConfig's transition scores, modelconfig.c:86-136) against the
fields the compiled reference built for the same sampled HMMs (tests/golden/
vit_golden.npz), and its length-dependent move cost against the reference's
table.  No GPU needed."""
import numpy as np

import cudahmmsearch_amd as ca
from conftest import vit_golden_profile


def test_conversion_bit_exact(vit_golden):
    z, man = vit_golden
    n = 0
    for i, m in enumerate(man["profiles"]):
        if m["seed"] is None:
            continue
        g = vit_golden_profile(z, man, i)
        v = ca.sample_profile(m["seed"], m["M"]).viterbi()
        assert np.array_equal(v.rwv, g["rwv"]), m["name"]
        assert np.array_equal(v.twv, g["twv"]), m["name"]
        assert (v.e_loop, v.e_move, v.base_w, v.ddbound_w) == (g["e_loop"], g["e_move"], g["base_w"], g["ddbound_w"])
        assert np.float32(v.scale_w).view(np.uint32) == m["scale_w_bits"]
        assert np.float32(v.nj).view(np.uint32) == m["nj_bits"]
        n += 1
    assert n >= 10


def test_move_cost_bit_exact(vit_golden):
    z, man = vit_golden
    for i, m in enumerate(man["profiles"]):
        g = vit_golden_profile(z, man, i)
        got = [ca.vit_move(g["scale_w"], g["nj"], int(L)) for L in z["move_L"]]
        assert got == list(z["moves"][i]), m["name"]
