"""Independent counts of an extraction, computed with torch from the case tables (no engine code):
crossing edges (= MC vertices = DMC quads), MC triangles per Bourke code and DMC patches per code
after the face-neighbour rule (src/cudualmc.cu:815-839), on the grid padded as the reference's
wrapper pads it (diso/__init__.py:52-54). Used where the oracle is too slow (large grids)."""
import torch
import torch.nn.functional as F

MC_CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
DMC_CORNERS = [(i & 1, (i >> 1) & 1, (i >> 2) & 1) for i in range(8)]


def padded_inside(sdf, iso=0.0):
    return F.pad(sdf, (1, 1, 1, 1, 1, 1), value=iso + 1.0) >= iso


def crossings(ins):
    return sum(int((ins.narrow(a, 0, ins.shape[a] - 1) != ins.narrow(a, 1, ins.shape[a] - 1)).sum()) for a in range(3))


def codes(ins, corners):
    X, Y, Z = (d - 1 for d in ins.shape)
    code = torch.zeros((X, Y, Z), dtype=torch.int64, device=ins.device)
    for i, (dx, dy, dz) in enumerate(corners):
        code |= ins[dx:dx + X, dy:dy + Y, dz:dz + Z].long() << i
    return code


def good_codes(code, pc):
    """The face-neighbour rule on a whole code grid: an ambiguous case is inverted when its
    neighbour in the tabulated direction exists and is ambiguous too."""
    d = pc[code]
    out = code.clone()
    dims = code.shape
    for comp in range(3):
        for step in (1, -1):
            sel = d == (2 * comp + (1 if step > 0 else 0))
            if not bool(sel.any()):
                continue
            idx = sel.nonzero(as_tuple=True)
            n = [t.clone() for t in idx]
            n[comp] = n[comp] + step
            ok = (n[comp] >= 0) & (n[comp] < dims[comp])
            nb = torch.full_like(idx[0], 255)
            nb[ok] = pc[code[n[0][ok], n[1][ok], n[2][ok]]]
            flip = ok & (nb != 255)
            out[idx[0][flip], idx[1][flip], idx[2][flip]] ^= 255
    return out


def expected_counts(sdf, tables, iso=0.0):
    """(crossing edges, MC triangles, DMC patches) of `sdf` [X,Y,Z]; `tables` are the expanded
    case tables as int64 device tensors (keys fmi, fpi, pc)."""
    ins = padded_inside(sdf, iso)
    nx = crossings(ins)
    ntri = int(((tables["fmi"][1:] - tables["fmi"][:-1]) // 3)[codes(ins, MC_CORNERS)].sum())
    dmc = good_codes(codes(ins, DMC_CORNERS), tables["pc"])
    npatch = int((tables["fpi"][1:] - tables["fpi"][:-1])[dmc].sum())
    return nx, ntri, npatch
