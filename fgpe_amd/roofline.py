"""Algorithmic HBM bytes of the fused time step (the byte model the roofline
fractions are quoted against).

Every kernel reads and writes its operands exactly once; an FFT pass moves a
spectral channel in and out once (2 * S_c bytes per point).  The spectral
channels are the ncomp complex components plus, when a k-space kernel is on,
the total density stored as a half spectrum: kx = 0..nx/2 of each x row, so
it moves a fraction h = (nx/2 + 1)/nx of a complex channel.  Per point, with
C = ncomp + h (density on) or ncomp (off):

  strided (y fwd or inv) / last (fwd + multiply + inv):  2 * C * S_c
      (+ h * S_r for the Vt table when a soft-core term is on, last only)
  xstage, stage s:  2 * C * S_c  plus, per component, the pointwise operands
      s=1: read P0, V;            write K, PS
      s=2,3: read PS, P0, V, K;   write K, PS
      s=4: read PS, P0, K, V;     write P0
with S_c the bulk complex size (16 FP64, 8 mixed), P0 always FP64 complex
(16), V real (S_c/2) and PS, K of the bulk complex size.
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s


def sizes(mixed: bool):
    sc = 8 if mixed else 16
    return sc, sc // 2, 16   # complex bulk, real V, FP64 state


def half_fraction(nx: int) -> float:
    """Share of a complex channel that a half-spectrum real channel moves."""
    return (nx // 2 + 1) / nx


def channels(ncomp: int, dens: bool, nx: int) -> float:
    return ncomp + (half_fraction(nx) if dens else 0.0)


def xstage_bytes(stage: int, mixed: bool, ncomp: int = 1, dens: bool = True, nx: int = 256) -> float:
    sc, sv, s0 = sizes(mixed)
    spec = 2 * channels(ncomp, dens, nx) * sc
    if stage == 1:
        pw = s0 + sv + sc + sc
    elif stage in (2, 3):
        # the mixed form rebuilds the state instead of reading the FP64 one: stage 2
        # from its input and K, stage 3 from its input, K and the kept stage-1 state
        pw = sc + ((0 if stage == 2 else sc) if mixed else s0) + sv + 2 * sc + sc
    else:
        pw = sc + 2 * s0 + sc + sv
    return spec + ncomp * pw


def pass_bytes(mixed: bool, ncomp: int = 1, dens: bool = True, nx: int = 256, table: bool = False) -> float:
    sc, sv, _ = sizes(mixed)
    return 2 * channels(ncomp, dens, nx) * sc + (half_fraction(nx) * sv if table else 0.0)


def step_bytes_per_point(mixed: bool, ncomp: int = 1, dens: bool = True, nx: int = 256, three_d: bool = True,
                         table: bool = False) -> float:
    """Bytes per grid point of one RK4 step (4 stages)."""
    plain = pass_bytes(mixed, ncomp, dens, nx)
    per_stage_passes = (2 * plain if three_d else 0) + pass_bytes(mixed, ncomp, dens, nx, table)
    return 4 * per_stage_passes + sum(xstage_bytes(s, mixed, ncomp, dens, nx) for s in (1, 2, 3, 4))


def adjoint_bytes_per_point(mixed: bool, nsteps: int, every: int, nx: int = 128) -> float:
    """Algorithmic bytes per grid point of one adjoint gradient of the
    one-component dipolar GPE (fgpe_adjoint_forward + fgpe_adjoint_backward):

      forward     nsteps RK4 steps, a checkpoint copy (32 B) per segment
      recompute   per segment: the checkpoint copied back and its x prologue,
                  every step's state copied to the segment buffer (32 B) and
                  the step run again, its RK stages writing their stage states
                  y1..y3 to per-step buffers instead of the stage buffer (the
                  same bytes)
      per step    four vector-Jacobian products: k_adj_pro (adjoint seed in, state
                  in, spectral channels out: the complex adjoint seed and the
                  real n and r as half spectra), the y / last / y passes on
                  those channels, k_adj_epi (channels, state, seed, trap,
                  lambda in and out, next seed out, dV accumulated).
    The adjoint vectors stay FP64 (16 B) in both forms; the state the
    products are taken at is the bulk stage state, or the FP64 step input."""
    sc, sv, s0 = sizes(mixed)
    h = half_fraction(nx)
    step = step_bytes_per_point(mixed, 1, True, nx)
    nseg = (nsteps + every - 1) // every
    fwd = nsteps * step + nseg * 32
    recompute = nseg * (32 + s0 + sc * (1 + h)) + nsteps * (32 + step)
    ch = (1 + 2 * h) * sc          # the VJP's spectral channels
    passes3 = 3 * 2 * ch
    vjp = 0.0
    for s in (3, 2, 1, 0):
        y = s0 if s == 0 else sc
        pro = (32 if s == 3 else 16) + y + ch
        epi = ch + y + 16 + sv + 16 + (0 if s == 3 else 16) + 16 + (16 if s > 0 else 0) + 16
        vjp += pro + passes3 + epi
    return fwd + recompute + nsteps * vjp
