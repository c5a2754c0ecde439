"""Oracle (test infrastructure only) for the frame output path: NumPy
restatement, operation by operation in float32, of

  * the mean of the FP64 accumulation buffer (this build's tfg_image_resolve;
    the reference averages a float32 host accumulator, main.cpp:1109-1111),
  * Imagen::acesFilmic (imagen.h:194-205),
  * the byte conversion of guardarPNG (imagen_utils.h:58-62):
    round(clamp(v, 0, 1) * 255) with halves rounded away from zero.

NaN handling follows C's fminf/fmaxf (the non-NaN operand wins).
"""
import numpy as np

F = np.float32


def mean(accum: np.ndarray, samples: int) -> np.ndarray:
    return (np.asarray(accum, np.float64) * (1.0 / float(samples))).astype(np.float32)


def aces_filmic(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, np.float32)
    a, b, c, d, e = F(2.51), F(0.03), F(2.43), F(0.59), F(0.14)
    with np.errstate(invalid="ignore", over="ignore"):
        x = x * F(0.6)
        num = x * (a * x + b)
        den = x * (c * x + d) + e
        y = num / den
    return np.fmin(np.fmax(y, F(0.0)), F(1.0))


def to_bytes(v: np.ndarray) -> np.ndarray:
    c = np.fmin(np.fmax(np.asarray(v, np.float32), F(0.0)), F(1.0)) * F(255.0)
    return np.floor(c + F(0.5)).astype(np.uint8)   # c >= 0: floor(c + 0.5) is round-half-away


def resolve(accum: np.ndarray, samples: int, tonemap: bool = True) -> np.ndarray:
    m = mean(accum, samples)
    return to_bytes(aces_filmic(m) if tonemap else m)
