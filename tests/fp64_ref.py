"""The unsplit FP64 MFMA kernel as the accuracy reference of the split forms.

The 16384^3 comparison of the split forms with FP64 runs the FP64 kernel without split-K (its 16384
output tiles fill the GPU). Reduced-size tests of the same comparison pin split-K off while they
compute the FP64 reference, so that the reference is that kernel's summation and not the partial sums
of 128-deep slices the tuned policy would choose for a few output tiles.
"""
import contextlib

from jlc_amd import native


@contextlib.contextmanager
def unsplit_fp64():
    old = native.tune_get("splitk_tiles_per_cu")
    native.tune_set("splitk_tiles_per_cu", 0)
    try:
        yield
    finally:
        native.tune_set("splitk_tiles_per_cu", old)
