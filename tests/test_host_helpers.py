"""Host-side helpers of the frame loop (no GPU)."""
import pytest

from galaxy_hip import track_halo_centers


def test_halo_centre_ema_moves_a_tenth_of_the_way():
    c = [[0.0, 0.0, 0.0], [10.0, -4.0, 2.0]]
    track_halo_centers(c, [[1.0, 2.0, -3.0], [0.0, 0.0, 2.0]])
    assert c[0] == pytest.approx([0.1, 0.2, -0.3])
    assert c[1] == pytest.approx([9.0, -3.6, 2.0])
    # only as many centres move as there are targets
    track_halo_centers(c, [[1.0, 2.0, -3.0]])
    assert c[1] == pytest.approx([9.0, -3.6, 2.0])
