"""Schedule table lookup (CPU)."""
from gpugo_amd import schedule as S


def test_schedule_lookup_caps_cycle():
    (wid, cyc), k = S.schedule_for("sepcos", 2)
    assert cyc <= 2 and wid >= 2 and k >= 1
    assert S.d_class(3) == 4 and S.d_class(17) == 32 and S.d_class(100) == 128
    assert S.fits((3, 4, 32), 10, 1 << 18) and not S.fits((3, 12, 1), 50, 1 << 18)
