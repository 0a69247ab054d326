"""Argument checks of the Python host mirror (tetwild_amd/ops.py), on the CPU.

The C-ABI cannot know the extent of a caller's buffer, so the mirror checks
every tensor it passes, outputs included, before the library is reached.  A
Context is stood up without a device here; a call that gets past the checks
reaches _bind(), which raises Reached instead of touching a device."""
import pytest
import torch

from tetwild_amd import _capi as C
from tetwild_amd.ops import Context


class Reached(Exception):
    """the call passed every argument check"""


@pytest.fixture
def ctx():
    c = Context.__new__(Context)
    c.device = torch.device("cpu")
    c.L = None
    c.h = None

    def bind():
        raise Reached()
    c._bind = bind
    return c


def inputs(n=5):
    tets = torch.zeros((n, 4), dtype=torch.int32)
    x = torch.zeros(8, dtype=torch.float64)
    return tets, x, x.clone(), x.clone()


def f64(*s):
    return torch.zeros(s, dtype=torch.float64)


def test_amips_out_checked(ctx):
    t, x, y, z = inputs()
    egh = C.TWG_AMIPS_EGH
    with pytest.raises(ValueError):
        ctx.amips_eval(t, x, y, z, egh, out=(f64(4), f64(5, 3), f64(5, 9)))      # short E
    with pytest.raises(ValueError):
        ctx.amips_eval(t, x, y, z, egh, out=(f64(5), f64(5, 3), None))          # H missing
    with pytest.raises(TypeError):
        ctx.amips_eval(t, x, y, z, egh, out=(f64(5), f64(5, 3), f64(5, 9).float()))
    with pytest.raises(ValueError):
        ctx.amips_eval(t, x, y, z, egh, out=(f64(5), f64(3, 5).t(), f64(5, 9)))  # not contiguous
    with pytest.raises(ValueError):
        ctx.amips_eval(t, x, y, z, egh, out=(f64(5), f64(5, 3)))                 # not 3 tensors
    with pytest.raises(ValueError):
        ctx.amips_launcher(t, x, y, z, egh, (f64(5), f64(5, 2), f64(5, 9)))
    # mode E only: G and H may be None
    with pytest.raises(Reached):
        ctx.amips_eval(t, x, y, z, C.TWG_AMIPS_E, out=(f64(5), None, None))
    # the packed Hessian mode takes H (n, 6), not (n, 9)
    sym = C.TWG_AMIPS_EGH_SYM
    with pytest.raises(ValueError):
        ctx.amips_eval(t, x, y, z, sym, out=(f64(5), f64(5, 3), f64(5, 9)))
    with pytest.raises(Reached):
        ctx.amips_eval(t, x, y, z, sym, out=(f64(5), f64(5, 3), f64(5, 6)))
    with pytest.raises(Reached):
        ctx.amips_eval(t, x, y, z, sym)


def test_amips_mixed_out_checked(ctx):
    t, x, y, z = inputs()
    g32 = torch.zeros((5, 3), dtype=torch.float32)
    h32 = torch.zeros((5, 9), dtype=torch.float32)
    fb = torch.zeros(5, dtype=torch.uint8)
    with pytest.raises(TypeError):
        ctx.amips_eval_mixed(t, x, y, z, out=(f64(5), f64(5, 3), h32, fb))        # G not float32
    with pytest.raises(ValueError):
        ctx.amips_eval_mixed(t, x, y, z, out=(f64(5), g32, h32, fb[:4]))          # short fallback
    with pytest.raises(Reached):
        ctx.amips_eval_mixed(t, x, y, z, out=(f64(5), g32, h32, fb))


def test_envelope_out_checked(ctx):
    tri = f64(6, 9)
    with pytest.raises(ValueError):
        ctx.envelope_faces(None, tri, 1e-6, 1e-3, out=torch.zeros(5, dtype=torch.uint8))
    with pytest.raises(TypeError):
        ctx.envelope_faces(None, tri, 1e-6, 1e-3, out=torch.zeros(6, dtype=torch.int32))
