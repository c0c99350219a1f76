"""A Tecplot FEBRICK flow file (IO_TECPLOTH::write, tecplot.cpp:16-177) reduced to
a form independent of node numbering and element order: one line per element with
its 8 corners' coordinates (connectivity order) and its U, V, W, exactly as
printed, the lines sorted; and the digest of that form."""
import hashlib
import re


def canonical(text):
    lines = text.splitlines()
    m = re.search(r"NODES=(\d+), ELEMENTS=(\d+)", lines[2])
    n_nodes, n_el = int(m.group(1)), int(m.group(2))
    vals = lines[3:3 + 3 * n_nodes + 3 * n_el]
    X, Y, Z = (vals[a * n_nodes:(a + 1) * n_nodes] for a in range(3))
    off = 3 * n_nodes
    U, V, W = (vals[off + a * n_el:off + (a + 1) * n_el] for a in range(3))
    conn = lines[3 + 3 * n_nodes + 3 * n_el:3 + 3 * n_nodes + 4 * n_el]
    rows = []
    for e in range(n_el):
        ids = [int(t) - 1 for t in conn[e].split()]
        rows.append(" ".join(f"{X[i]},{Y[i]},{Z[i]}" for i in ids) + f" | {U[e]} {V[e]} {W[e]}")
    rows.sort()
    return n_nodes, n_el, rows


def digest(text):
    n_nodes, n_el, rows = canonical(text)
    return {"nodes": n_nodes, "elements": n_el, "sha256": hashlib.sha256("\n".join(rows).encode()).hexdigest()}
