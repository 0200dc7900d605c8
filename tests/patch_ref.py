"""Python restatement of Get_PointPatch (PPR/Get_PointPatch.f90:83-500) for uniform
rectangular meshes, used to check the library's patch builder and to feed the CPU oracle
in CPU-only tests. Node ids are 1-based like the reference's point_patch."""
import numpy as np


def element_and_node_classes(nnx, nny, xs, ys, circles):
    """node_index (-2 inside a circle, -1 vacuum) and element_index (>0 interface
    counter, -1 vacuum, -2 object) by node classification."""
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    inside = np.zeros_like(X, dtype=bool)
    for (cx, cy, r) in circles:
        inside |= (X - cx) ** 2 + (Y - cy) ** 2 < r * r
    node_index = np.where(inside, -2, -1).ravel().astype(np.int32)   # node (i,j) -> i*nny + j
    nex, ney = nnx - 1, nny - 1
    element_index = np.zeros(nex * ney, dtype=np.int32)
    k = 0
    for ci in range(nex):
        for cj in range(ney):
            e = cj + ci * ney
            kin = int(inside[ci, cj]) + int(inside[ci + 1, cj]) + int(inside[ci + 1, cj + 1]) + int(inside[ci, cj + 1])
            if kin == 4:
                element_index[e] = -2
            elif kin == 0:
                element_index[e] = -1
            else:
                k += 1
                element_index[e] = k
    return node_index, element_index


def point_patches(nnx, nny, element_index):
    nex, ney = nnx - 1, nny - 1
    nnodes = nnx * nny

    def el(ci, cj):
        ci = min(max(ci, 0), nex - 1)
        cj = min(max(cj, 0), ney - 1)
        return cj + ci * ney + 1

    def nodes_of(e1):
        e = e1 - 1
        ci, cj = divmod(e, ney)
        n1 = ci * nny + cj
        return {n1 + 1, n1 + nny + 1, n1 + nny + 2, n1 + 2}

    pelem, ptype = [], []
    for i in range(nnx):
        for j in range(nny):
            xl, xr, yb, yt = i == 0, i == nnx - 1, j == 0, j == nny - 1
            ex = {1: el(i - 1, j - 1), 2: el(i, j - 1), 3: el(i, j), 4: el(i - 1, j)}
            if xl and yb:
                els, t = [1], 2
            elif xl and yt:
                els, t = [el(0, ney - 1)], 2
            elif xr and yb:
                els, t = [el(nex - 1, 0)], 2
            elif xr and yt:
                els, t = [el(nex - 1, ney - 1)], 2
            elif xl:
                els, t = [ex[2], ex[3]], 2
            elif yt:
                els, t = [ex[1], ex[2]], 2
            elif xr:
                els, t = [ex[1], ex[4]], 2
            elif yb:
                els, t = [ex[4], ex[3]], 2
            else:
                e4 = [ex[1], ex[2], ex[3], ex[4]]
                idx = [element_index[e - 1] for e in e4]
                if all(v == -2 for v in idx) or all(v == -1 for v in idx):
                    els, t = e4, 1
                else:
                    els, t = [e for e in e4 if element_index[e - 1] < 0], 3
            pelem.append(sorted(set(els)))
            ptype.append(t)
    patch = []
    for n in range(nnodes):
        s = set()
        for e in pelem[n]:
            s |= nodes_of(e)
        patch.append(sorted(s))
    big = [list(p) for p in patch]
    for n in range(nnodes):
        if 0 < len(patch[n]) < 6:
            s = set()
            for k in big[n]:
                s |= set(big[k - 1])
            patch[n] = sorted(s)
    for n in range(nnodes):
        if ptype[n] in (2, 3) and len(patch[n]) < 12:
            s = set()
            for k in big[n]:
                s |= set(big[k - 1])
            if s:
                patch[n] = sorted(s)
    pp = np.zeros((nnodes, 50), dtype=np.int32)   # point_patch(nnodes, 50)
    for n in range(nnodes):
        pp[n, 0] = len(patch[n])
        pp[n, 1:1 + len(patch[n])] = patch[n]
    return pp.ravel(order="F"), ptype
