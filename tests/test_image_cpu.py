"""Frame output on the host side: the oracle's tone curve and byte conversion
pinned by closed forms, and the PNG writer by a decode round trip."""
import numpy as np

from oracle import image_oracle as O
from tfg_amd import image


def test_aces_closed_forms():
    # imagen.h:194-205: f(0) = 0; f grows to a / c = 1.033 > 1, so large
    # inputs clamp to 1; compare a mid value with the double-precision curve
    assert O.aces_filmic(np.float32(0.0)) == 0.0
    assert O.aces_filmic(np.float32(1e6)) == 1.0
    x = np.linspace(0.0, 8.0, 4001, dtype=np.float32)
    xd = x.astype(np.float64) * 0.6
    ref = np.clip(xd * (2.51 * xd + 0.03) / (xd * (2.43 * xd + 0.59) + 0.14), 0.0, 1.0)
    np.testing.assert_allclose(O.aces_filmic(x), ref, atol=2e-7)
    assert np.all(np.diff(O.aces_filmic(x)) >= 0)


def test_bytes_round_half_away_and_clamp():
    v = np.array([0.5, -1.0, 2.0, 0.0, 1.0, np.nan, 1.0 / 255.0, 0.5 / 255.0], np.float32)
    b = O.to_bytes(v)
    # 0.5 * 255 = 127.5 exactly -> 128 (std::round), NaN -> 0 (fmaxf)
    np.testing.assert_array_equal(b, [128, 0, 255, 0, 255, 0, 1, 1])


def test_mean_uses_the_frame_count():
    acc = np.array([[3.0, 6.0, 9.0]])
    np.testing.assert_array_equal(O.mean(acc, 3), np.array([[1.0, 2.0, 3.0]], np.float32))


def test_png_round_trip():
    rng = np.random.default_rng(1)
    for w, h in ((1, 1), (7, 12), (64, 3)):
        px = rng.integers(0, 256, (w * h, 3), dtype=np.uint8)
        data = image.png_bytes(px, w, h)
        assert data[:8] == b"\x89PNG\r\n\x1a\n"
        w2, h2, back = image.read_png_rgb8(data)
        assert (w2, h2) == (w, h)
        np.testing.assert_array_equal(back, px)
