"""Write brute-force golden vectors with networkx (run where networkx is
installed; the tests only read the committed fixtures).

For small seeded graphs and queries, networkx's VF2
``DiGraphMatcher(G, Q, node_match, edge_match).subgraph_monomorphisms_iter()``
lists every injective map of query vertices to data vertices that keeps every
query edge (with its label) -- exactly the embeddings GFSM counts on symmetric
simple inputs.  Each fixture stores the data graph, the query and the sorted
embedding rows (column c = query vertex c).  With ``vlabel`` the node labels
must also match.

Output: tests/golden/nx_fixtures.json
"""
import json
import os
import random

import networkx as nx
from networkx.algorithms import isomorphism as iso

HERE = os.path.dirname(os.path.abspath(__file__))


def rand_graph(rng, n, p, nvl, nel):
    g = nx.Graph()
    for v in range(n):
        g.add_node(v, label=rng.randint(1, nvl))
    for a in range(n):
        for b in range(a + 1, n):
            if rng.random() < p:
                g.add_edge(a, b, label=rng.randint(1, nel))
    return g


def walk_query(rng, g, q, dense):
    nodes = [v for v in g.nodes if g.degree(v) > 0]
    for _ in range(1000):
        start = rng.choice(nodes)
        chosen, tree, cur, stall = [start], [], start, 0
        while len(chosen) < q and stall < 100:
            nb = rng.choice(list(g.neighbors(cur)))
            if nb not in chosen:
                tree.append((chosen.index(cur), len(chosen)))
                chosen.append(nb)
                stall = 0
            else:
                stall += 1
            cur = nb
        if len(chosen) == q:
            break
    else:
        return None
    qg = nx.Graph()
    for i, v in enumerate(chosen):
        qg.add_node(i, label=g.nodes[v]["label"])
    if dense:
        for i in range(q):
            for j in range(i + 1, q):
                if g.has_edge(chosen[i], chosen[j]):
                    qg.add_edge(i, j, label=g.edges[chosen[i], chosen[j]]["label"])
    else:
        for i, j in tree:
            qg.add_edge(i, j, label=g.edges[chosen[i], chosen[j]]["label"])
    return qg


def entries(g):
    s, d, l = [], [], []
    for a, b, data in g.edges(data=True):
        s += [a, b]
        d += [b, a]
        l += [data["label"], data["label"]]
    return s, d, l


def embeddings(g, qg, vlabel):
    G = g.to_directed()
    Q = qg.to_directed()
    em = iso.categorical_edge_match("label", None)
    nm = iso.categorical_node_match("label", None) if vlabel else None
    m = iso.DiGraphMatcher(G, Q, node_match=nm, edge_match=em)
    rows = []
    for mp in m.subgraph_monomorphisms_iter():  # data vertex -> query vertex
        row = [0] * qg.number_of_nodes()
        for dv, qv in mp.items():
            row[qv] = dv
        rows.append(row)
    rows.sort()
    return rows


def main():
    rng = random.Random(20261202)
    fixtures = []
    specs = [
        # (n, p, nvl, nel, q, dense, vlabel)
        (12, 0.35, 1, 1, 3, True, False), (12, 0.35, 1, 1, 4, True, False), (14, 0.3, 1, 2, 4, False, False),
        (16, 0.25, 2, 1, 4, True, True), (16, 0.25, 2, 2, 5, False, True), (18, 0.22, 3, 1, 5, True, True),
        (18, 0.22, 1, 3, 5, False, False), (20, 0.2, 2, 2, 4, True, True), (20, 0.25, 1, 1, 5, True, False),
        (22, 0.18, 2, 1, 6, False, True), (10, 0.6, 1, 1, 4, True, False), (10, 0.6, 2, 1, 5, True, True),
        (24, 0.15, 1, 1, 6, False, False), (24, 0.2, 3, 2, 4, False, True), (15, 0.4, 2, 2, 3, True, True),
        (8, 1.0, 1, 1, 4, True, False),
    ]
    for k, (n, p, nvl, nel, q, dense, vlabel) in enumerate(specs):
        g = rand_graph(rng, n, p, nvl, nel)
        qg = walk_query(rng, g, q, dense)
        if qg is None or qg.number_of_edges() == 0:
            continue
        ds, dd, dl = entries(g)
        qs, qd, ql = entries(qg)
        fixtures.append({
            "name": f"nx{k:02d}_n{n}_q{q}_{'dense' if dense else 'sparse'}_{'vlabel' if vlabel else 'ref'}",
            "mode": 1 if vlabel else 0,
            "data": {"V": n, "vlabel": [g.nodes[v]["label"] for v in range(n)], "src": ds, "dst": dd, "elabel": dl},
            "query": {"V": q, "vlabel": [qg.nodes[v]["label"] for v in range(q)], "src": qs, "dst": qd,
                      "elabel": ql},
            "rows": embeddings(g, qg, vlabel),
        })
    with open(os.path.join(HERE, "nx_fixtures.json"), "w") as f:
        json.dump({"generator": f"networkx {nx.__version__} DiGraphMatcher.subgraph_monomorphisms_iter",
                   "fixtures": fixtures}, f, separators=(",", ":"))
    print(len(fixtures), "fixtures;", [len(x["rows"]) for x in fixtures])


if __name__ == "__main__":
    main()
