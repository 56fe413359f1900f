"""Write the golden fixtures under tests/golden/ (run from the repo root).

The reference ships no golden vectors for this path, so the fixtures are
computed from independent formulations:

* topology.json -- boundary edges / nodes of the meshes the reference's own
  tests use (2-triangle quad of test2D/test_2Dfem_elastic.py:17-18, the
  16-segment circle of flatworld/femesher.py:132-149, the 100x100 grid of
  test2D/test_2Dfem_elastic10000.py:19-32), from the dict restatement of
  flatworld/mesh.py:53-71 in oracle/oracle.py;
* materials.json -- PK2 stresses at the inputs of
  test2D/test_2Dmaterial_functions.py:19-72 (E=1e6, nu=0.3), obtained as
  S = dW/dE of the textbook strain energies by central differences in
  float64:
    St.Venant-Kirchhoff  W = lam/2 (tr E)^2 + mu tr(E^2)
    compressible NH      W = mu/2 (tr C - 2) - mu ln J + lam/2 (ln J)^2
  (plane strain, C = F^T F = I + 2E, J = sqrt(det C)).
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import boundary_edges_dict  # noqa: E402


def quad_mesh():
    conn = np.array([[0, 1, 3], [0, 3, 2]], dtype=np.int32)
    coords = np.array([[0.5, 0.5], [0.7, 0.5], [0.5, 0.7], [0.7, 0.7]], dtype=np.float32)
    return conn, coords


def circle_mesh(origin=(0.5, 0.5), radius=0.1):
    x, y = origin
    verts = [[x + radius * np.cos(2.0 * np.pi * i / 16), y + radius * np.sin(2.0 * np.pi * i / 16)] for i in range(16)]
    verts.append([x, y])
    conn = [[i, (i + 1) % 16, 16] for i in range(16)]
    return np.array(conn, dtype=np.int32), np.array(verts, dtype=np.float32)


def grid100_mesh():
    n = 100
    dx = 0.6 / n
    conn = np.zeros((2 * n * n, 3), dtype=np.int32)
    coords = np.zeros(((n + 1) * (n + 1), 2), dtype=np.float32)
    k = 0
    for i in range(n + 1):
        for j in range(n + 1):
            coords[k] = np.array([dx * i, dx * j])
            k += 1
    e = 0
    for i in range(n):
        for j in range(n):
            conn[e] = [i * (n + 1) + j, i * (n + 1) + j + 1, (i + 1) * (n + 1) + j]
            conn[e + 1] = [i * (n + 1) + j + 1, (i + 1) * (n + 1) + j + 1, (i + 1) * (n + 1) + j]
            e += 2
    return conn, coords + 0.3


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def w_stvk(Evoigt, lam, mu):
    e00, e11, e01 = Evoigt
    tr = e00 + e11
    return 0.5 * lam * tr * tr + mu * (e00 * e00 + e11 * e11 + 2 * e01 * e01)


def w_nh(Evoigt, lam, mu):
    e00, e11, e01 = Evoigt
    C = np.array([[1 + 2 * e00, 2 * e01], [2 * e01, 1 + 2 * e11]])
    lnJ = 0.5 * np.log(np.linalg.det(C))
    return 0.5 * mu * (np.trace(C) - 2.0) - mu * lnJ + 0.5 * lam * lnJ * lnJ


def stress_from_energy(W, F, lam, mu, h=1e-7):
    F = np.asarray(F, dtype=np.float64)
    Cm = F.T @ F
    Ev = np.array([0.5 * (Cm[0, 0] - 1), 0.5 * (Cm[1, 1] - 1), 0.5 * Cm[0, 1]])
    S = np.zeros(3)
    # dW/dE_xx, dW/dE_yy; the shear derivative counts E01 and E10 together (factor 1/2)
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        g = (W(Ev + d, lam, mu) - W(Ev - d, lam, mu)) / (2 * h)
        S[k] = g if k < 2 else 0.5 * g
    return S


def main():
    topo = {}
    for name, (conn, coords) in {"quad": quad_mesh(), "circle": circle_mesh(), "grid100": grid100_mesh()}.items():
        be = boundary_edges_dict(conn)
        topo[name] = {
            "n_nodes": int(coords.shape[0]),
            "n_elems": int(conn.shape[0]),
            "boundary_edges": be.tolist(),
            "boundary_nodes": sorted(set(int(v) for v in be.reshape(-1))),
        }
    with open(os.path.join(HERE, "topology.json"), "w") as f:
        json.dump(topo, f)

    E, nu = 1.0e6, 0.3
    lam, mu = lame(E, nu)
    cases = {
        "identity": [[1.0, 0.0], [0.0, 1.0]],
        "uniaxial_1.01": [[1.01, 0.0], [0.0, 1.0]],
        "shear_0.01": [[1.0, 0.01], [0.0, 1.0]],
        "stretch_1.1": [[1.1, 0.0], [0.0, 1.0]],
        "general": [[1.03, -0.02], [0.015, 0.97]],
    }
    mats = {"E": E, "nu": nu, "cases": {}}
    for name, F in cases.items():
        mats["cases"][name] = {
            "F": F,
            "elastic_S": stress_from_energy(w_stvk, F, lam, mu).tolist(),
            "neohookean_S": stress_from_energy(w_nh, F, lam, mu).tolist(),
        }
    with open(os.path.join(HERE, "materials.json"), "w") as f:
        json.dump(mats, f, indent=1)
    print("wrote topology.json, materials.json")


if __name__ == "__main__":
    main()
