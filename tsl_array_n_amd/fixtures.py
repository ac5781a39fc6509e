"""Pressure-system fixtures in the reference's on-disk format
(export_system.mjs:65-126): a directory holding meta.json and one
little-endian float32 .bin per array, first index fastest --
b, pressure, dirichletMask, ventMask (0/1 as floats) and, when the operator
has face weights, weightU/V/W.

export_system() writes the assembled system of a GridPressureSolver3 (after
assemble() or project()); load_system() reads a fixture written by it or by
the reference's capture script; solve_system() solves a loaded system on the
GPU with the library's operator, V-cycle and PCG.
"""
import json
import os

import numpy as np


def _face_sizes(res):
    res = list(res)
    out = []
    for a in range(3):
        s = list(res)
        s[a] += 1
        out.append(s)
    return out


def export_system(directory, solver, diagnostics=None, source="tsl_array_n_amd", frame=0, multigrid=None,
                  tolerance=None, max_iterations=None, grid_spacing=None):
    """Write the solver's assembled system (b, current pressure, masks,
    operator face weights) as a fixture directory; returns the meta dict."""
    os.makedirs(directory, exist_ok=True)
    res = [int(x) for x in solver.resolution]
    arrays = {
        "b": solver.b.to_array().astype("<f4"),
        "pressure": solver.pressure.to_array().astype("<f4"),
        "dirichletMask": (solver.dirichlet_mask.to_array().astype("<f4") if solver.dirichlet_mask is not None
                          else None),
        "ventMask": solver.vent_mask.to_array().astype("<f4") if solver.vent_mask is not None else None,
    }
    ws = solver.weights
    for a, name in enumerate(("weightU", "weightV", "weightW")):
        arrays[name] = ws[a].to_array().astype("<f4") if ws is not None else None
    written = []
    for name, values in arrays.items():
        path = os.path.join(directory, name + ".bin")
        if values is None:
            if os.path.exists(path):
                os.remove(path)
            continue
        values.tofile(path)
        written.append(name)
    fs = _face_sizes(res)
    meta = {
        "source": source,
        "frame": frame,
        "resolution": res,
        "gridSpacing": list(grid_spacing) if grid_spacing is not None else [1, 1, 1],
        "dataSizeU": fs[0], "dataSizeV": fs[1], "dataSizeW": fs[2],
        "diagnostics": dict(diagnostics or getattr(solver, "diagnostics", {}) or {}),
        "multigrid": dict(multigrid or {}),
        "tolerance": tolerance,
        "maxIterations": max_iterations,
        "arrays": written,
    }
    with open(os.path.join(directory, "meta.json"), "w") as f:
        json.dump(meta, f, indent="\t")
        f.write("\n")
    return meta


def load_system(directory):
    """Read a fixture directory: meta, shape, spacing, b, pressure, the masks
    as uint8 (> 0.5) and the face weights (None where absent)."""
    with open(os.path.join(directory, "meta.json")) as f:
        meta = json.load(f)
    shape = tuple(int(x) for x in meta["resolution"])
    cells = int(np.prod(shape))

    def rd(name, count):
        path = os.path.join(directory, name + ".bin")
        if not os.path.exists(path):
            return None
        v = np.fromfile(path, dtype="<f4")
        if v.size != count:
            raise ValueError("%s: %d floats, expected %d" % (path, v.size, count))
        return v.astype(np.float32)

    fs = _face_sizes(shape)
    ws = [rd(n, int(np.prod(s))) for n, s in zip(("weightU", "weightV", "weightW"), fs)]
    dm = rd("dirichletMask", cells)
    vm = rd("ventMask", cells)
    return {
        "meta": meta,
        "shape": shape,
        "spacing": tuple(float(h) for h in meta.get("gridSpacing", [1, 1, 1])),
        "b": rd("b", cells),
        "pressure": rd("pressure", cells),
        "mask": (dm > 0.5).astype(np.uint8) if dm is not None else None,
        "vent": (vm > 0.5).astype(np.uint8) if vm is not None else None,
        "weights": ws if all(w is not None for w in ws) else None,
    }


def solve_system(F, system, levels=4, tol=1e-6, maxiter=None, x0=None, storage="f32"):
    """Solve a loaded fixture on the GPU (createLaplacianOperator +
    createMultigridPreconditioner + PCG); returns (pressure, state)."""
    shape, spacing = system["shape"], system["spacing"]
    mask = F.array_n("u8", shape, system["mask"]) if system["mask"] is not None else None
    vent = F.array_n("u8", shape, system["vent"]) if system["vent"] is not None else None
    weights = None
    if system["weights"] is not None:
        weights = [F.array_n("float", F.face_shape(shape, a), system["weights"][a]) for a in range(3)]
    op = F.create_laplacian_operator(shape, spacing, mask, weights)
    mg = F.create_multigrid_preconditioner(shape, spacing, dirichlet_mask=mask, coarse_dirichlet_mask=vent,
                                           face_weights=weights, levels=levels, storage=storage)
    b = F.array_n("float", shape, system["b"])
    x = F.array_n("float", shape, system["pressure"] if x0 is None else x0)
    pcg = F.create_pcg_solver(op, "multigrid", b, x, multigrid=mg)
    cap = maxiter or int(np.ceil(1.5 * max(shape)))
    pcg.solve(tol, cap, relative_tolerance=True)
    return x.to_array(), pcg.state
