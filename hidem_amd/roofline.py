"""Compulsory HBM bytes of the fused step, for roofline figures (bench.py,
hidem_amd.batched).  MI355X HBM3E peak: 8 TB/s."""
from __future__ import annotations

from . import _abi as A

HBM_PEAK_GBS = 8000.0


def step_bytes(n, n_neigh_total, n_links, prec, exact=True):
    """Compulsory HBM bytes of one fused step launch (k_step: link break of
    the previous step, contact + beam force gather, integration).

    Per particle: coordinates, rotations, Diff (12 x 8), radius (8), type (1),
    neighbour and link counts (2 x 4), and for the exact rebuild test the
    coordinates at the last rebuild (3 x 8) read; coordinates, rotations and
    Diff (12 x 8) and the force (6 x 8 FP64 / 6 x 4 mixed) written.  Per
    neighbour-list entry: its 4 B index.  Per link side: CT0..5 read and
    written (2 x 6 x 8 FP64 / 2 x 6 x 4 mixed), link word and other endpoint
    (2 x 4).  Per link: initialLen and maxLoad (2 x 8) for the break test.
    Partner state is gathered from particles already counted once (cache
    reuse); the rare writes of a breaking link are not counted."""
    fo = 8 if prec == A.HDM_PREC_FP64 else 4
    per_particle = 96 + 8 + 1 + 8 + (24 if exact else 0) + 96 + 6 * fo
    per_side = 12 * fo + 8
    return n * per_particle + 4 * n_neigh_total + 2 * n_links * per_side + 16 * n_links


def reference_list_entries(n_list, n_links, exec_info):
    """Entries of the reference's neighbour list: the mixed form's list leaves
    out link partners (their contacts run in the link loop), which the
    reference lists, two entries per link."""
    return n_list + (2 * n_links if exec_info["merged_contacts"] else 0)
