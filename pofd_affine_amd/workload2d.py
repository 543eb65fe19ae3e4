"""Device-path context for a synthetic two-band workload (pofd_affine_amd.
synth2d): the host preparation the reference's calcLikeDouble does when
reading its spec file, then the C ABI set-up."""
from . import _lib, synth2d


def build_context2d(inp, n, theta, nedge=256, binned=False, edge_integ=True,
                    sigma=(synth2d.SIGMA1, synth2d.SIGMA2), device=0,
                    spectral_cut=None):
    ctx = _lib.Context2(device)
    if spectral_cut is not None:
        ctx.set_spectral_cut(spectral_cut)
    kp, sp, op = synth2d.positions()
    ctx.configure2d(n, kp, sp, op, nedge=nedge, edge_integ=edge_integ)
    bm = _lib.DoubleBeam(inp["beam1"], inp["beam2"], inp["pixsize"],
                         inp["minbeamval"], inp["nbeamhist"])
    b = ctx.add_beamset2d(bm)
    if binned:
        ctx.add_field2d(b, inp["d1"], inp["d2"], sigma[0], sigma[1],
                        bins=inp["bins"], bininfo=inp["bininfo"])
    else:
        ctx.add_field2d(b, inp["d1"], inp["d2"], sigma[0], sigma[1])
    ctx.prepare2d(theta)
    return ctx
