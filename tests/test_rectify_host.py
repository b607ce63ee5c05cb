"""Host-side parity of the rectify maps' set-up (no device needed): the
product's inverse(Knew R) is the oracle's bit for bit, and a singular
rectification is rejected before any launch."""
import numpy as np
import pytest

import pyoracle as O
import vxsgm
from test_oracle_rectify import rot


@pytest.mark.parametrize("seed", range(5))
def test_inverse_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    p = vxsgm.RectifyParams(fx=rng.uniform(300, 900), fy=rng.uniform(300, 900),
                            cx=rng.uniform(100, 700), cy=rng.uniform(50, 400),
                            R=tuple(rot(*rng.uniform(-0.1, 0.1, 3)).ravel()),
                            new_fx=rng.uniform(300, 900), new_fy=rng.uniform(300, 900),
                            new_cx=rng.uniform(100, 700), new_cy=rng.uniform(50, 400))
    _, _, m = O.rectify_maps(p.as_dict(), 1, 1)
    assert (vxsgm.rectify_inverse(p).view(np.uint32) == m.view(np.uint32)).all()


def test_singular_rejected_on_host():
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.rectify_inverse(vxsgm.RectifyParams(fx=500.0, new_fx=0.0))
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.rectify_maps_enqueue(vxsgm.RectifyParams(), 0, 4, 1, 1)
