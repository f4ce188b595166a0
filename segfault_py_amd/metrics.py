"""Work and flop accounting for the J/K build (quartet counts per class and
the algorithmic flop model used for roofline fractions).

Flop model per primitive shell quartet of class (la lb|lc ld), with
L = la+lb+lc+ld, nr = L//2 + 1 Rys nodes and n(l) = (l+1)(l+2)/2:

    F = 40                                   # P, Q, rho, X, prefactors
      + 60 nr                                # nodes and weights
      + 3 nr 4 (la+lb+1)(lc+ld+1)            # 2-D recurrence
      + 3 nr 2 (la+1)(lb+1)(lc+1)(ld+1)      # horizontal transfer
      + n(la) n(lb) n(lc) n(ld) (4 nr + 2)   # Ix Iy Iz w products + contraction

plus 12 flops per unique function quartet for the J/K digestion.
"""
from __future__ import annotations

import numpy as np

from .engine import NUM_QUARTET_CLASSES, PAIR_CLASSES, quartet_class_ls


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def flops_per_prim_quartet(la, lb, lc, ld) -> int:
    L = la + lb + lc + ld
    nr = L // 2 + 1
    return (40 + 60 * nr + 3 * nr * 4 * (la + lb + 1) * (lc + ld + 1)
            + 3 * nr * 2 * (la + 1) * (lb + 1) * (lc + 1) * (ld + 1)
            + ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) * (4 * nr + 2))


# Executed flops per primitive quartet of the point-charge form of far
# quartets (eri_one_node_far): P-Q 3, |P-Q|^2 5, 1/|P-Q| (rsqrt seed + two
# Newton steps) 13, S_b S_k / R 2, screening test 1, accumulation 1; (ps|ss)
# adds c = ip/(2R^2) 3, P-A 3, three components 3 x (fma + mul + add) less
# the one accumulation.  The algorithmic count (class_work's flops) stays
# the Rys figure per primitive quartet, the fixed unit of work.
FAR_FLOPS = {(0, 0, 0, 0): 25, (1, 0, 0, 0): 42}


def class_work(basis, mol=None, evaluated=None, far=None) -> list[dict]:
    """Per quartet class: shell / primitive / unique-function quartet counts
    (all surviving, and the FP64 part) and algorithmic flops.  Primitive
    quartets count the primitive pairs kept by primitive-pair screening;
    ``evaluated`` ([55, 2] from Basis.prim_quartet_counts(0)), when given,
    replaces the all-quartet count by the primitive quartets the kernels
    evaluate after primitive-quartet screening, so flops count only work
    done; ``far`` ([55, 2] from Basis.far_prim_quartet_counts(0)) records
    the part of them evaluated in the point-charge form (``primq_far``;
    FAR_FLOPS executed flops each, the algorithmic flops unchanged)."""
    out = []
    pairs = {}
    for pc in range(len(PAIR_CLASSES)):
        sa, sb, _, npp = basis.pair_list(pc, with_nprim=True)
        pairs[pc] = (sa, sb, npp.astype(np.int64))
    for c in range(NUM_QUARTET_CLASSES):
        la, lb, lc, ld = quartet_class_ls(c)
        pcb = la * (la + 1) // 2 + lb
        pck = lc * (lc + 1) // 2 + ld
        n_all, n64 = basis.quartet_list(c)
        rec = {"class": c, "ls": (la, lb, lc, ld), "shellq": 0, "shellq64": 0, "primq": 0,
               "primq64": 0, "funcq": 0, "funcq64": 0, "flops": 0.0, "flops64": 0.0}
        if len(n_all) and len(pairs[pck][0]):
            sa, sb, npb = pairs[pcb]
            ka, kb, npk = pairs[pck]
            nA, nB, nC, nD = ncart(la), ncart(lb), ncart(lc), ncart(ld)
            pab = np.where(sa == sb, nA * (nA + 1) // 2, nA * nB).astype(np.int64)
            pcd = np.where(ka == kb, nC * (nC + 1) // 2, nC * nD).astype(np.int64)
            cp = np.concatenate([[0], np.cumsum(npk)])
            cf = np.concatenate([[0], np.cumsum(pcd)])
            same = pcb == pck
            for key, n in (("", n_all), ("64", n64)):
                n = n.astype(np.int64)
                rec["shellq" + key] = int(n.sum())
                rec["primq" + key] = int((npb * cp[n]).sum())
                fq = pab * cf[n]
                if same:
                    # the diagonal quartet (i|i) counts pab(pab+1)/2, not pab^2
                    idx = np.arange(len(n))
                    diag = n > idx
                    fq = fq - np.where(diag, pab * pab - pab * (pab + 1) // 2, 0)
                rec["funcq" + key] = int(fq.sum())
            if evaluated is not None:
                rec["primq"] = int(evaluated[c][0])
            fpq = flops_per_prim_quartet(la, lb, lc, ld)
            # far (ss|ss) / (ps|ss) primitive quartets are the same units of
            # work in the point-charge form (fewer executed flops, FAR_FLOPS)
            rec["primq_far"] = int(far[c][0]) if far is not None else 0
            rec["flops"] = float(rec["primq"] * fpq + 12 * rec["funcq"])
            rec["flops64"] = float(rec["primq64"] * fpq + 12 * rec["funcq64"])
        out.append(rec)
    return out
