"""Algorithmic HBM bytes of the fused time step (the byte model the roofline
fractions are quoted against).

Every kernel reads and writes its operands exactly once; an FFT pass moves a
spectral channel in and out once (2 * S_c bytes per point).  Per point:

  strided (y fwd or inv) / last (fwd + multiply + inv):  2 * nch * S_c
      (+ S_r for the Vt table when a soft-core term is on, last only)
  xstage, stage s (1 component):  2 * nch * S_c  plus the pointwise operands
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


def xstage_bytes(stage: int, mixed: bool, nch: int) -> int:
    sc, sv, s0 = sizes(mixed)
    spec = 2 * nch * sc
    if stage == 1:
        pw = s0 + sv + sc + sc
    elif stage in (2, 3):
        pw = sc + s0 + sv + 2 * sc + sc
    else:
        pw = sc + 2 * s0 + sc + sv
    return spec + pw


def pass_bytes(mixed: bool, nch: int, table: bool = False) -> int:
    sc, sv, _ = sizes(mixed)
    return 2 * nch * sc + (sv if table else 0)


def step_bytes_per_point(mixed: bool, nch: int, three_d: bool = True, table: bool = False) -> int:
    """Bytes per grid point of one RK4 step (4 stages)."""
    per_stage_passes = (2 * pass_bytes(mixed, nch) if three_d else 0) + pass_bytes(mixed, nch, table)
    return 4 * per_stage_passes + sum(xstage_bytes(s, mixed, nch) for s in (1, 2, 3, 4))
