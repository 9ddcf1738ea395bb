"""Writes tests/golden/known_answers.json: the known answers the reference holds for this path.

The reference has no automated tests; its known answers are printed literals in tutorial
programs.  They are transcribed here as data (inputs and expected outputs), each with the
file:line it comes from.  The CRS graph rows of the 2x1x1 Cube3D are derived by hand from
the first-appearance rule of ZeroMatrix_as_CRS_FEMDomain (FEMDomainClass.f90:10680-10734):

  nodes (6 per layer): layer 1 ids 1..6 at (x,y) = (0,0),(1,0),(2,0),(0,1),(1,1),(2,1);
  layer 2 ids 7..12 above them.  Elements (MeshClass.f90:6796-6806, :7065-7070):
  e1 = [1,2,5,4, 7,8,11,10], e2 = [2,3,6,5, 8,9,12,11].
  A node in e1 only lists e1's nodes; a node in e2 only lists e2's nodes; a node in both
  lists e1's nodes, then the nodes of e2 not yet seen: 3,6,9,12.

Run:  python tests/golden/make_golden.py
"""
import json
from pathlib import Path

E1 = [1, 2, 5, 4, 7, 8, 11, 10]
E2 = [2, 3, 6, 5, 8, 9, 12, 11]
BOTH = E1 + [3, 6, 9, 12]
ROWS_2x1x1 = {1: E1, 2: BOTH, 3: E2, 4: E1, 5: BOTH, 6: E2, 7: E1, 8: BOTH, 9: E2, 10: E1, 11: BOTH, 12: E2}


def graph_2x1x1():
    row_ptr = [1]
    col = []
    for node in range(1, 13):
        col += ROWS_2x1x1[node]
        row_ptr.append(len(col) + 1)
    return {"source": "FEMDomainClass.f90:10665-10838 applied by hand to Cube3D 2x1x1",
            "nn": 12, "dof": 1, "elemnod": [E1, E2], "row_ptr": row_ptr, "col_idx": col}


def main():
    data = {
        "graph_1x1x1_node1": {
            "source": "FEMDomainClass.f90:10693-10696 applied by hand to Cube3D 1x1x1 (hex local order)",
            "row": [1, 2, 4, 3, 5, 6, 8, 7]},
        "graph_2x1x1_dof1": graph_2x1x1(),
        "cube3d_2x1x1_elemnod": {"source": "MeshClass.f90:6796-6806 (2-D) + :7065-7070 (layers)",
                                 "elemnod": [E1, E2]},
        "systems": [
            {"name": "ex0007_easySolver", "source": "Tutorial/std/ex0007_easySolver.f90:15-30",
             "A": [[1, 0, 0, 0], [0, 2, -1, 0], [0, -1, 2, 0], [0, 0, 0, 1]],
             "b": [0.5, 0.5, 0.2, 0.2], "x": [0.5, 0.4, 0.3, 0.2]},
            {"name": "ex0004_linearsolvers", "source": "Tutorial/std/ex0004_linearsolvers.f90:19-36",
             "A": [[2, -1, 0, 0], [-1, 2, -1, 0], [0, 2, 3, 1], [0, 0, 1, 2]],
             "b": [1, 2, 3, 4], "x": None},
            {"name": "PBiCGSTAB_ILU", "source": "Tutorial/std/PBiCGSTAB_ILU.f90:10-31",
             "A": [[5, 6, 7], [10, 20, 23], [15, 50, 67]], "b": [2, 3, 4], "x": [0.54, -0.35, 0.20]},
            {"name": "MPI_BiCGSTAB_global", "source": "Tutorial/fem/MPI_BiCGSTAB_example.f90:63-92 "
             "(printed global matrix; RHS(7)=1 as set at :38-40)",
             "A": [[1, 0, 0, 0, 0, 0, 0], [0, 2, -1, 0, 0, 0, 0], [0, -1, 2, -1, 0, 0, 0],
                   [0, 0, -1, 2, -1, 0, 0], [0, 0, 0, -1, 2, -1, 0], [0, 0, 0, 0, -1, 2, -1],
                   [0, 0, 0, 0, 0, -1, 1]],
             "b": [4, 4, 0, 0, 0, 0, 1], "x": [4, 5, 6, 7, 8, 9, 10]},
        ],
        "bar_in_gravity": {"source": "Tutorial/fem/benchmark_bar_in_gravity_field.f90:12-67",
                           "mesh": [10, 10, 100], "E": 2.17e7, "nu": 0.0, "rho": 2.4, "g": 9.8,
                           "theory_kPa": -234.024, "simulated_kPa": -234.026},
    }
    out = Path(__file__).with_name("known_answers.json")
    out.write_text(json.dumps(data, indent=1))
    print("wrote", out)


if __name__ == "__main__":
    main()
