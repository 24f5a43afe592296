"""Generate mc_tables.h: the Lewiner marching-cubes case and tiling tables.

Source: scikit-image's skimage/measure/_marching_cubes_lewiner_luts.py, itself
generated from LookUpTable.h of Lewiner et al., "Efficient implementation of
Marching Cubes' cases with topological guarantees" (JGT 8(2), 2003).  The
tables are data that must match the reference's marching cubes
(skimage.measure.marching_cubes, method='lewiner', i3d/meshing.py:296-298)
entry for entry, so they are transcribed, not re-derived.  The file stores
each table as (shape, base64 of int8 values); this script decodes them into C
arrays.  Usage:
    python gen_mc_tables.py /path/to/_marching_cubes_lewiner_luts.py > mc_tables.h
"""
import ast
import base64
import sys

import numpy as np

ORDER = ["CASES", "TILING1", "TILING2", "TILING3_1", "TILING3_2", "TILING4_1", "TILING4_2",
         "TILING5", "TILING6_1_1", "TILING6_1_2", "TILING6_2", "TILING7_1", "TILING7_2",
         "TILING7_3", "TILING7_4_1", "TILING7_4_2", "TILING8", "TILING9", "TILING10_1_1",
         "TILING10_1_1_", "TILING10_1_2", "TILING10_2", "TILING10_2_", "TILING11",
         "TILING12_1_1", "TILING12_1_1_", "TILING12_1_2", "TILING12_2", "TILING12_2_",
         "TILING13_1", "TILING13_1_", "TILING13_2", "TILING13_2_", "TILING13_3", "TILING13_3_",
         "TILING13_4", "TILING13_5_1", "TILING13_5_2", "TILING14", "TEST3", "TEST4", "TEST6",
         "TEST7", "TEST10", "TEST12", "TEST13", "SUBCONFIG13"]


def load(path):
    tree = ast.parse(open(path).read())
    tables = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and isinstance(node.value, ast.Tuple):
            name = node.targets[0].id
            shape = ast.literal_eval(node.value.elts[0])
            text = ast.literal_eval(node.value.elts[1])
            arr = np.frombuffer(base64.decodebytes(text.encode()), dtype=np.int8).reshape(shape)
            tables[name] = arr
    return tables


def c_array(name, arr):
    dims = "".join(f"[{d}]" for d in arr.shape)
    flat = ",".join(str(int(v)) for v in arr.reshape(-1))
    return f"MC_TABLE int8_t MC_{name}{dims} = {{{flat}}};"


def main():
    tables = load(sys.argv[1])
    out = ["// Generated by gen_mc_tables.py from the Lewiner marching-cubes tables (see there).",
           "// Include after defining MC_TABLE (e.g. `static const` or `__device__ static const`).",
           "#pragma once", "#include <stdint.h>"]
    for name in ORDER:
        out.append(c_array(name, tables[name]))
    print("\n".join(out))


if __name__ == "__main__":
    main()
