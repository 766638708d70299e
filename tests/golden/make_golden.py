"""Writes tests/golden/toy_golden.json.

Inputs are the two toy CSR graphs that the reference itself ships
(Basic/Graph/graphFileList.hpp:510-570, `getToyGraph("toy_0" / "toy_1")`),
copied here as data.  Expected outputs are computed WITHOUT the project's
oracle, by libraries and brute force:

  BFS  networkx.single_source_shortest_path_length, cross-checked by a
       pure-Python Bellman-Ford with unit weights;
  SSSP networkx.single_source_dijkstra_path_length (zero-weight edges are
       honoured), cross-checked by pure-Python Bellman-Ford;
  CC   (directed graph, the reference's push semantics) min id over all
       vertices that reach v, from networkx.ancestors; and on the
       symmetrized graph, min id per networkx weakly connected component;
  PR   dense float64 power iteration with numpy, d = 0.85, 20 iterations,
       dangling mass dropped.

Unreached vertices take the reference sentinel 999 999 999
(Basic/Graph/basic_def.hpp:18).  Run: python tests/golden/make_golden.py
"""
import json
import os

import networkx as nx
import numpy as np

UNREACHED = 999_999_999

TOYS = {
    "toy_0": {
        "offsets": [0, 3, 6, 6, 8, 8, 11, 15, 16, 21, 22, 24, 26],
        "dest": [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5],
        "weight": [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5],
    },
    "toy_1": {
        "offsets": [0, 3, 6, 6, 8, 8, 11, 16, 17, 21, 24, 26, 28],
        "dest": [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 10, 5, 0, 4, 9, 11, 5, 6, 8, 2, 8, 2, 5],
        "weight": [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5, 2, 5],
    },
}
ROOTS = [0, 3, 6]


def edges(t):
    off, dst, w = t["offsets"], t["dest"], t["weight"]
    for u in range(len(off) - 1):
        for e in range(off[u], off[u + 1]):
            yield u, dst[e], w[e]


def bellman_ford(n, es, root, unit):
    d = [None] * n
    d[root] = 0
    for _ in range(n):
        for u, v, w in es:
            if d[u] is not None:
                c = d[u] + (1 if unit else w)
                if d[v] is None or c < d[v]:
                    d[v] = c
    return [UNREACHED if x is None else x for x in d]


def main():
    out = {"_source": "inputs: reference getToyGraph (Basic/Graph/graphFileList.hpp:510-570); "
                      "outputs: networkx / numpy / brute force, see make_golden.py",
           "unreached": UNREACHED, "pr_damping": 0.85, "pr_iterations": 20, "graphs": {}}
    for name, t in TOYS.items():
        n = len(t["offsets"]) - 1
        es = list(edges(t))
        g = nx.DiGraph()
        g.add_nodes_from(range(n))
        for u, v, w in es:
            g.add_edge(u, v, weight=w)
        rec = {"V": n, "E": len(es), **t, "bfs": {}, "sssp": {}}
        for r in ROOTS:
            lv = nx.single_source_shortest_path_length(g, r)
            bfs = [lv.get(v, UNREACHED) for v in range(n)]
            assert bfs == bellman_ford(n, es, r, True), (name, r)
            dd = nx.single_source_dijkstra_path_length(g, r, weight="weight")
            sssp = [dd.get(v, UNREACHED) for v in range(n)]
            assert sssp == bellman_ford(n, es, r, False), (name, r)
            rec["bfs"][str(r)] = bfs
            rec["sssp"][str(r)] = sssp
        rec["cc_directed"] = [min({v} | nx.ancestors(g, v)) for v in range(n)]
        ug = g.to_undirected()
        lab = [0] * n
        for comp in nx.connected_components(ug):
            m = min(comp)
            for v in comp:
                lab[v] = m
        rec["cc_symmetric"] = lab
        # Dense power iteration: M[v, u] = 1/outdeg(u) for u -> v.
        M = np.zeros((n, n))
        for u in range(n):
            deg = t["offsets"][u + 1] - t["offsets"][u]
            for e in range(t["offsets"][u], t["offsets"][u + 1]):
                M[t["dest"][e], u] += 1.0 / deg
        pr = np.full(n, 1.0 / n)
        for _ in range(20):
            pr = 0.15 / n + 0.85 * (M @ pr)
        rec["pr"] = [float(x) for x in pr]
        out["graphs"][name] = rec
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "toy_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
