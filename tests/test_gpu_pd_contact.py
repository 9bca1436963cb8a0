"""Contact between the particles of a PD specimen on the GPU (pdk_set_contact:
Solver with a contact model, CabanaPD_Solver.hpp:134-151, 332-404;
NormalRepulsion, force/CabanaPD_Contact.hpp:44-80) against the oracle, bit
for bit: two PMB fracture blocks fly at each other, touch only through the
contact model (no bond joins them), shatter and rebound."""
import numpy as np
import pytest

from cabanapd_amd import configs
from tests.parity import pair, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_colliding_blocks(precision):
    say(f"case: colliding PMB blocks 2 x 8^3 with NormalRepulsion contact, {precision}")
    w = configs.colliding_blocks(steps=200, precision=precision)
    g, o = pair(w)
    assert g.contact_pairs() == o.contact_pairs() == 0
    for k in range(5):
        g.step(40)
        o.step(40)
        pg, po = g.contact_pairs(), o.contact_pairs()
        ig, io = g.bonds()[2], o.bonds()[2]
        eu = float(np.max(np.abs(g.get("u") - o.get("u"))) / np.max(np.abs(o.get("u"))))
        say(f"  step {(k + 1) * 40}: contact pairs {pg}/{po}, broken {(ig == 0).sum()}/{(io == 0).sum()}, "
            f"u rel diff {eu:.3e}")
        if precision == "fp64":
            assert pg == po
            assert np.array_equal(ig, io)
            for fld in ("u", "v", "f"):
                assert np.array_equal(g.get(fld), o.get(fld)), fld
    v = g.get("v")
    n = v.shape[0] // 2
    # momentum exchange through contact only: the blocks rebound
    assert v[:n, 0].mean() < 0 < v[n:, 0].mean()
    g.close()
    o.close()
