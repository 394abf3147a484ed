"""Child process of tests/test_sanitizers.py: runs the FP64 oracle built under
AddressSanitizer + UndefinedBehaviorSanitizer on inputs read from an .npz and
writes every output to another .npz.  Imports numpy and the oracle's ctypes
wrapper only (the process preloads gcc's libasan).

    python tests/_asan_oracle_child.py <sanitized .so> <inputs.npz> <outputs.npz>
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SHADER_KEYS = ("normal", "view_dir", "albedo", "specular_albedo", "roughness", "envmap",
               "light2obj_rotmat", "visibility", "face_vertex_list", "nearest_triangle_id",
               "barycentric_coord", "grad_shading", "grad_diffuse_shading")


def run(inputs: dict) -> dict:
    out = {}
    d = {k: inputs[k] for k in SHADER_KEYS if k in inputs}
    for v in ("diffuse", "specular", "clamp", "v2", "vis"):
        for diff, spec in ((1, 1), (1, 0), (0, 1)):
            sh, ds = O.forward(v, d, diff, spec)
            out[f"{v}{diff}{spec}/shading"] = sh
            if ds is not None:
                out[f"{v}{diff}{spec}/diffuse_shading"] = ds
            g = O.backward(v, d, d["grad_shading"], d.get("grad_diffuse_shading"), diff, spec,
                           roughness_abs=True)
            for k, x in g.items():
                out[f"{v}{diff}{spec}/grad_{k}"] = x
    m = {k[5:]: inputs[k] for k in inputs if k.startswith("mesh/")}
    tri, bary = O.nearest_mesh_points_forward(m["adjacency_head"], m["adjacency_list"],
                                              m["face_vertex_list"], m["vertex_pos"],
                                              m["query_pos"], m["idxs"])
    out["mesh/tri"], out["mesh/bary"] = tri, bary
    out["mesh/grad_query"] = O.nearest_mesh_points_backward(
        m["face_vertex_list"], m["vertex_pos"], m["query_pos"], tri, m["grad_bary"])
    out["uv/fwd"] = O.uv_of_triangle_forward(inputs["uv/triangles"], inputs["uv/query"])
    out["uv/bwd"] = O.uv_of_triangle_backward(inputs["uv/triangles"], inputs["uv/query"],
                                              inputs["uv/grad"])
    out["bits"] = O.extract_bitfield(inputs["visibility"], 40, 3, 17)
    out["env/sample"] = O.sample_envmap(inputs["env/query"], inputs["env/map"])
    out["env/blur"] = O.blur_envmap(inputs["env/map"], 4, 8, 96, 0.3)
    return out


if __name__ == "__main__":
    O._LIB_PATH = sys.argv[1]
    with np.load(sys.argv[2]) as f:
        inputs = {k: f[k] for k in f.files}
    np.savez(sys.argv[3], **run(inputs))
    print("oracle under ASan+UBSan:", O.lib()._name, flush=True)
