"""Why per-path parity bars are below 100%, checked path by path.

The HIP tracer shades in FP32 (hardware rcp/sqrt/sin/cos); the oracle
shades in FP64.  Both draw the same PCG32 numbers, so a path can only leave
the oracle's result through a discrete decision that rounding moves: which
triangle a ray hits near an edge, whether a shadow ray grazes an occluder, a
Russian-roulette or lobe threshold, total internal reflection.  Such a path
is ill-conditioned in the oracle itself: displacing each of its rays by a
tiny relative amount (the oracle's conditioning probe,
Renderer.pixels_perturbed) changes its result beyond the tolerance.

explain() splits the paths that disagree with the oracle into those the
probe shows to be ill-conditioned (explained by rounding) and the rest, and
gives the base rate of ill-conditioned paths over the whole image.  A
systematic error (a wrong branch, a wrong sign) affects well-conditioned
paths as often as any others, so its disagreements would be explained at
about the base rate, far below the bar the tests set.
"""
import numpy as np

TOL = 1e-3          # per-channel relative tolerance of the per-path bars
DELTA = 1e-5        # relative displacement: an upper bound of FP32 vs FP64 drift over a path
SEEDS = 4


def agree(a, b, tol=TOL):
    return np.all(np.abs(np.asarray(a, np.float64) - b) <= tol * (1.0 + np.abs(b)), axis=1)


def ill_conditioned(renderer, pix, ref, frame, max_depth, rr_depth, delta=DELTA, seeds=SEEDS):
    """Per pixel index in `pix` (row-major over the renderer's image): True if
    any of `seeds` displaced oracle renders leaves the tolerance band around
    the unperturbed oracle result `ref` (rows matching `pix`)."""
    pix = np.asarray(pix, np.int64)
    out = np.zeros(pix.size, bool)
    if pix.size == 0:
        return out
    xy = np.stack([pix % renderer.width, pix // renderer.width], 1)
    for k in range(seeds):
        p = renderer.pixels_perturbed(xy, frame, max_depth, rr_depth, delta, 1000 + 17 * k)
        out |= ~agree(p, ref)
    return out


def explain(renderer, got, ref, frame, max_depth, rr_depth, base_sample=2048, seed=0):
    """(n_disagree, n_explained, base_rate) for one frame: got and ref are
    (n_pixels, 3) images of the same frame (HIP path, oracle)."""
    ok = agree(got, ref)
    bad = np.nonzero(~ok)[0]
    ill = ill_conditioned(renderer, bad, ref[bad], frame, max_depth, rr_depth)
    rng = np.random.default_rng(seed)
    n = ref.shape[0]
    sample = rng.choice(n, size=min(base_sample, n), replace=False)
    base = ill_conditioned(renderer, sample, ref[sample], frame, max_depth, rr_depth).mean()
    return int(bad.size), int(ill.sum()), float(base)
