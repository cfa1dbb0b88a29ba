"""CPU restatement of the reference's phase-defect line analysis.

Test infrastructure only (the checker of rdcu_amd.lines): numpy and
scipy.sparse, following makelines.py
  * wloop      (makelines.py:19-24)   winding of the phase around each plaquette;
  * findlines  (makelines.py:26-57)   cells next to a defect plaquette, joined into
                                      lines where cells are at periodic L1
                                      distance 1; each line as its edges in
                                      physical coordinates;
  * correlate  (makelines.py:112-144) minimum periodic L1 distance between the
                                      cells of each line of one frame and each
                                      line of the next.
networkx's connected components are replaced by scipy.sparse.csgraph on the
same adjacency; lines are ordered by their smallest cell index (the order
networkx yields them in) and each line's edges are listed once, (u, v) with
u < v in flat cell index, sorted.
"""
import numpy as np
from scipy.sparse import coo_array
from scipy.sparse.csgraph import connected_components


def wloop(th, axis1=0, axis2=1):
    """Winding number (integer) of the phase th around the plaquette spanned
    by axis1 and axis2 at each grid point (makelines.py:19-24)."""
    def wrap(d):
        return np.mod(d + np.pi, 2 * np.pi) - np.pi
    r1 = np.roll(th, -1, axis=axis1)
    r12 = np.roll(r1, -1, axis=axis2)
    r2 = np.roll(th, -1, axis=axis2)
    return np.round((wrap(r1 - th) + wrap(r12 - r1) + wrap(r2 - r12) + wrap(th - r2)) /
                    (2 * np.pi))


def line_cells(th, Ns):
    """Sorted flat indices of the cells sharing a face with a defect plaquette
    (makelines.py:28-42)."""
    Ns = np.asarray(Ns)
    sets = []
    for a1, a2, shift in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        c = np.array(np.nonzero(wloop(th, a1, a2)))
        sets.append(c)
        s = c.copy()
        s[shift] = np.mod(s[shift] - 1 + Ns[shift], Ns[shift])
        sets.append(s)
    allc = np.concatenate(sets, axis=1)
    if allc.shape[1] == 0:
        return np.zeros(0, dtype=np.int64)
    return np.unique(np.ravel_multi_index(tuple(allc), tuple(Ns)))


def _periodic_l1(a, b, Ns):
    d = a[:, None, :] - b[None, :, :]
    for ax in range(3):
        d[:, :, ax] = np.mod(d[:, :, ax] + Ns[ax] // 2, Ns[ax]) - Ns[ax] // 2
    return np.abs(d).sum(axis=-1)


def findlines(th, Ns, Ls, chunk=4096):
    """Lines (connected components) as lists of edges, each edge two cell
    positions in physical units (makelines.py:26-57): array [nedges, 2, 3]."""
    Ns = np.asarray(Ns)
    Ls = np.asarray(Ls, dtype=float)
    flat = line_cells(th, Ns)
    cells = np.array(np.unravel_index(flat, tuple(Ns))).T  # [nnz, 3]
    nnz = len(flat)
    rows, cols = [], []
    for n0 in range(0, nnz, chunk):
        dist = _periodic_l1(cells[n0:n0 + chunk], cells, Ns)
        r, c = np.nonzero(dist == 1)
        rows.append(r + n0)
        cols.append(c)
    r = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
    c = np.concatenate(cols) if cols else np.zeros(0, dtype=np.int64)
    adj = coo_array((np.ones(len(r)), (r, c)), shape=(nnz, nnz))
    ncomp, lab = connected_components(adj, directed=False) if nnz else (0, np.zeros(0, int))
    # components in order of their smallest node (node order = flat index order)
    first = {}
    for i, l in enumerate(lab):
        first.setdefault(l, i)
    order = sorted(first, key=lambda l: first[l])
    keep = r < c
    eu, ev = r[keep], c[keep]
    scale = Ls / (Ns - 1)
    lines = []
    for l in order:
        sel = (lab[eu] == l)
        e = np.stack([eu[sel], ev[sel]], axis=1)
        e = e[np.lexsort((e[:, 1], e[:, 0]))]
        pos = cells[e] * scale  # [nedges, 2, 3]
        lines.append(pos.reshape(-1, 2, 3))
    return lines


def correlate(prev, cur, Ns, Ls):
    """diffs[j, i] = minimum periodic L1 distance between the cells of line i
    of the previous frame and line j of the current one (makelines.py:112-144:
    previous-frame positions are rounded back to cells, current-frame ones
    truncated, as there)."""
    Ns = np.asarray(Ns)
    scale = np.asarray(Ls, dtype=float) / (Ns - 1)

    def cells_of(lines, rnd):
        out = []
        for l in lines:
            if len(l) == 0:  # the reference's concatenate fails here too
                raise ValueError("a line without edges (an isolated cell) has no cells")
            q = l / scale
            arr = np.concatenate(np.round(q).astype(int) if rnd else q.astype(int), axis=0)
            flat = np.unique(np.ravel_multi_index(tuple(arr.T), tuple(Ns)))
            out.append(np.array(np.unravel_index(flat, tuple(Ns))).T)
        return out

    c1, c2 = cells_of(prev, True), cells_of(cur, False)
    diffs = np.zeros((len(cur), len(prev)), dtype=int)
    for i, a in enumerate(c1):
        for j, b in enumerate(c2):
            diffs[j, i] = _periodic_l1(a, b, Ns).min()
    return diffs
