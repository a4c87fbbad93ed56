"""Edge-list preprocessing, the counterpart of preprocess/conversion.py.

    python -m thegraphpartitioning_amd.convert <file.{mtx,mmio,tsv,txt,tsv_1}> [out_dir]

Reads the edge list (conversion.py:30-58), drops self loops, makes every edge
bidirectional once and removes isolated vertices (66-69), then writes row.txt,
column.txt, weight.txt, output_output.csr and temp.graph (71-110) into out_dir
(default: the working directory), ready for `gpu_part <dir> <n> <m> <k>`.
The work is done by the native library (tgp_convert_edges,
tgp_write_conversion_files).
"""
import os
import sys
import time


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print(__doc__.strip().splitlines()[2].strip(), file=sys.stderr)
        return 2
    import thegraphpartitioning_amd as tgp
    path = argv[0]
    out_dir = argv[1] if len(argv) > 1 else "."
    ext = os.path.splitext(path)[1]
    print(f"File extension: {ext}")
    if ext not in (".mtx", ".mmio", ".tsv", ".txt", ".tsv_1"):
        print(f"Unsupported file type: {ext}")
        print("Supported: .tsv, .txt, .mtx, .mmio")
        return 1
    t0 = time.time()
    g = tgp.convert_edges(path)
    print(f"Graph Loaded. Time: {round(time.time() - t0, 4)}s")
    print(f"Nodes: {g.n}, Edges: {g.m}")
    print(f"CSR size: row_ptr={g.n + 1}, col_idx={g.m}")
    os.makedirs(out_dir, exist_ok=True)
    tgp.write_conversion_files(g, out_dir)
    return 0


if __name__ == "__main__":
    sys.exit(main())
