"""Reader for ASCII PLY point clouds as PCL writes them (test helper): the
vertex element's float properties by name; other elements (faces, camera)
are skipped.  The reference reads such files with pcl::io::loadPLYFile
(mesh/PoissonReconstruction.cpp:66-95)."""
import numpy as np


def read_ascii_ply(path):
    """Returns {property name: float32 array} for the vertex element."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError("not a PLY file")
        fmt = f.readline().split()
        if fmt[:2] != [b"format", b"ascii"]:
            raise ValueError("only ASCII PLY is read here")
        elements = []  # (name, count, [property names])
        while True:
            ln = f.readline()
            if not ln:
                raise ValueError("header without end_header")
            tok = ln.split()
            if not tok or tok[0] in (b"comment", b"obj_info"):
                continue
            if tok[0] == b"end_header":
                break
            if tok[0] == b"element":
                elements.append((tok[1].decode(), int(tok[2]), []))
            elif tok[0] == b"property":
                if tok[1] == b"list":
                    elements[-1][2].append(("list", tok[-1].decode()))
                else:
                    elements[-1][2].append((tok[1].decode(), tok[-1].decode()))
        out = {}
        for name, count, props in elements:
            rows = [f.readline() for _ in range(count)]
            if name != "vertex":
                continue
            if any(t == "list" for t, _ in props):
                raise ValueError("list properties in the vertex element")
            a = np.array([r.split() for r in rows], dtype=np.float64).reshape(count, len(props))
            for k, (_, pname) in enumerate(props):
                out[pname] = a[:, k].astype(np.float32)
        if "x" not in out:
            raise ValueError("no vertex element")
        return out
