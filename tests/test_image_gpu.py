"""GPU frame output (tfg_image_resolve) against the oracle restatement: the
bytes of the tone-mapped mean are bit-identical, for random accumulations,
special values and a rendered frame; the result is written as a PNG and
read back."""
import numpy as np
import pytest

from oracle import image_oracle as O
from tfg_amd import image, pt, scenes

pytestmark = pytest.mark.gpu


def test_resolve_matches_oracle(torch_cuda, tmp_path):
    torch = torch_cuda
    rng = np.random.default_rng(4)
    acc = rng.exponential(2.0, (5000, 3)) * rng.integers(1, 64, (5000, 1))
    acc[:8] = [[0, 0, 0], [-1, -2, -3], [1e30, 1e30, 1e30], [np.nan, 1, 2], [np.inf, 0, 0], [0.5, 0.5, 0.5],
               [63.75, 63.75, 63.75], [1e-30, 0, 0]]
    d = torch.from_numpy(acc).cuda()
    for samples in (1, 7, 64):
        for tonemap in (True, False):
            print(f"samples={samples} tonemap={tonemap}", flush=True)
            b, m = image.resolve(torch, d, samples, tonemap=tonemap, mean=True)
            torch.cuda.synchronize()
            gm, om = m.cpu().numpy(), O.mean(acc, samples)
            np.testing.assert_array_equal(gm, om)   # NaN where the oracle has NaN
            fin = np.isfinite(om)
            np.testing.assert_array_equal(gm[fin].view(np.uint32), om[fin].view(np.uint32))
            np.testing.assert_array_equal(b.cpu().numpy(), O.resolve(acc, samples, tonemap))


def test_rendered_frame_to_png(torch_cuda, tmp_path):
    torch = torch_cuda
    sc = scenes.cornell(level=3)
    W, H = 64, 48
    ds = pt.DeviceScene(sc, torch)
    tr = pt.PathTracer(W, H, 5, 2, mode=0, use_graph=False)
    tr.set_scene(ds)
    tr.set_camera(sc.camera.frame(W, H))
    img = torch.zeros((H * W, 3), dtype=torch.float32, device="cuda")
    acc = torch.zeros((H * W, 3), dtype=torch.float64, device="cuda")
    for f in range(16):
        tr.render(f, img, acc)
    b = image.resolve(torch, acc, 16)
    torch.cuda.synchronize()
    px = b.cpu().numpy()
    np.testing.assert_array_equal(px, O.resolve(acc.cpu().numpy(), 16))
    assert px.max() > 0
    path = tmp_path / "frame.png"
    image.write_png(path, px, W, H)
    w, h, back = image.read_png_rgb8(path.read_bytes())
    assert (w, h) == (W, H)
    np.testing.assert_array_equal(back, px)
