"""GPU parity of the point-light animation (bounce_point_lights.comp:33-73) through the C ABI, bit-exact against
the CPU oracle: NaNs (a light on a wall with a zero direction component there, 0 / 0 in the shader) compare equal
as NaNs, every other value bit for bit."""
import numpy as np
import pytest

from gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

LO, HI = (-50.0, -50.0, -50.0), (50.0, 50.0, 50.0)


def _lights(n, seed):
    # C2-like lights: positions U[-60, 60]^3 (about a fifth start outside the box and are clamped), unit
    # directions; a few exactly-zero direction components, a few lights exactly on a wall
    rng = np.random.default_rng(seed)
    pos = np.concatenate([rng.uniform(-60, 60, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dirs = np.concatenate([d, np.zeros((n, 1))], 1).astype(np.float32)
    if n >= 64:
        k = rng.choice(n, n // 16, replace=False)
        dirs[k, rng.integers(0, 3, k.size)] = 0.0
        w = rng.choice(n, n // 32, replace=False)
        pos[w, rng.integers(0, 3, w.size)] = np.float32(50.0)
    return pos, dirs


def _same(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(np.where(nan, 0, a).view(np.uint32), np.where(nan, 0, b).view(np.uint32))


@pytest.mark.parametrize("n", [1, 255, 4096, 100003])
@pytest.mark.parametrize("speed,dt", [(10.0, 1.0 / 60.0), (35.0, 0.5), (400.0, 1.0)])
def test_bounce_point_lights_bit_exact(vctx, oracle, n, speed, dt):
    # (400, 1): hundreds of units per call, so many lights use all 32 segments
    pos, dirs = _lights(n, 1000 + n)
    dp, dd = to_dev(pos), to_dev(dirs)
    want_p, want_d = pos, dirs
    for frame in range(3):  # consecutive frames of the animation, in place
        vctx.bounce_point_lights(dp, dd, n, LO, HI, speed, dt)
        want_p, want_d = oracle.bounce_point_lights(want_p, want_d, LO, HI, speed, dt)
        got_p, got_d = from_dev(dp, np.float32, (n, 4)), from_dev(dd, np.float32, (n, 4))
        assert _same(got_p, want_p) and _same(got_d, want_d), (n, speed, dt, frame)


def test_bounce_point_lights_no_lights(vctx):
    buf = empty_dev(64, fill=0x5A)
    vctx.bounce_point_lights(buf, buf, 0, LO, HI, 1.0, 1.0)
    assert (from_dev(buf, np.uint8) == 0x5A).all()
