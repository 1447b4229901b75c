"""The ViterbiFilter oracle (oracle/vit_oracle.c) against the golden vectors
written from the compiled reference (tests/golden/make_vit_golden.py):
every case's status and score bits, the length-dependent N/C/J move cost,
and the two facts the restatement relies on -- vf_conversion pads every
array with -32768 past the model and leaves the N/C/J loop costs at 0."""
import numpy as np

from conftest import vit_golden_cases, vit_golden_profile
from oracle import Oracle


def test_padding_and_loops(vit_golden):
    z, man = vit_golden
    for m in man["profiles"]:
        assert m["padding_is_min"], m["name"]
        assert m["loops_zero"], m["name"]


def test_move_table(vit_golden):
    z, man = vit_golden
    O = Oracle()
    for i in range(len(man["profiles"])):
        v = vit_golden_profile(z, man, i)
        got = [O.vit_move(v, int(L)) for L in z["move_L"]]
        assert got == list(z["moves"][i]), v["name"]


def test_scores_bit_exact(vit_golden):
    z, man = vit_golden
    O = Oracle()
    nbad = 0
    for i in range(len(man["profiles"])):
        v = vit_golden_profile(z, man, i)
        for c, p, x in vit_golden_cases(z, i):
            st, sc = O.vit(v, x)
            ok = st == z["case_status"][c] and np.float32(sc).view(np.uint32) == z["case_bits"][c]
            nbad += 0 if ok else 1
    assert nbad == 0
    assert (z["case_status"] == 16).sum() > 0 and (z["case_status"] == 0).sum() > 0


def test_plain_viterbi_bound():
    """What the mathematics fixes, independent of the reference: on a model
    whose only path is M1..MM with every transition probability 1 (so all
    transition costs are 0 except the entry), the filter's score for the
    consensus is the sum of the match costs plus the N/C/J and entry costs,
    and one mismatch lowers it by exactly the difference of the two costs."""
    O = Oracle()
    M = 5
    rwv = np.full((29, M), -1000, np.int16)
    for k in range(M):
        rwv[k, k] = 400 + 10 * k                     # residue k is the consensus at node k+1
    twv = np.zeros((8, M), np.int16)
    twv[0, :] = -2000                                # B->Mk entry
    twv[2, :] = twv[3, :] = -32768                   # no I->M, D->M
    twv[4:, :] = -32768                              # no M->D, M->I, I->I, D->D
    twv[1, 0] = -32768                               # no M0->M1
    v = {"rwv": rwv, "twv": twv, "e_loop": -500, "e_move": -500, "base_w": 12000, "ddbound_w": -32768,
         "scale_w": 500.0 / np.log(2), "nj": 1.0}
    x = np.arange(M, dtype=np.uint8)
    st, sc = O.vit(v, x)
    move = O.vit_move(v, M)
    xC = 12000 + move - 2000 + sum(400 + 10 * k for k in range(M)) - 500
    expect = np.float32((np.float32(xC) + np.float32(move) - np.float32(12000)) / np.float32(v["scale_w"]))
    assert st == 0 and np.isclose(sc, float(expect) - 3.0, rtol=0, atol=1e-4)
    y = x.copy()
    y[2] = 7                                         # mismatch at node 3: -1000 instead of 420
    st2, sc2 = O.vit(v, y)
    assert st2 == 0
    # single-segment paths after the mismatch: the full run with it (-2000 + 400 + 410
    # - 1000 + 430 + 440 = -1320), nodes 1-2 (-1190) or nodes 4-5 (-1130); two segments
    # joined through J pay another entry and cannot win.  The best, -1130, is 1230 units
    # below the consensus run's -2000 + 2100 = 100.
    assert np.isclose((sc - sc2) * v["scale_w"], 1230, atol=0.01)
