"""Write the small unlabeled query graphs used by the closed-form tests.

The shapes are those the reference benchmarks with (edge, 3-path, triangle,
4-path, 3-star, 4-cycle, 4-star "plus", 5-path; build/query/*.g): undirected,
one vertex label and one edge label, both directions listed.
"""
import os

SHAPES = {
    "road2": (2, [(0, 1)]),
    "road3": (3, [(0, 1), (1, 2)]),
    "triangle": (3, [(0, 1), (1, 2), (2, 0)]),
    "road4": (4, [(0, 1), (1, 2), (2, 3)]),
    "roadY": (4, [(0, 1), (0, 2), (0, 3)]),
    "square": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
    "plus": (5, [(1, 0), (1, 2), (1, 3), (1, 4)]),
    "road5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
}


def tve(n, edges):
    lines = ["t # 0", f"{n} {2 * len(edges)} 1 1"]
    lines += [f"v {i} 1" for i in range(n)]
    for a, b in edges:
        lines += [f"e {a} {b} 1", f"e {b} {a} 1"]
    lines.append("t # -1")
    return "\n".join(lines)


def main(out_dir=os.path.join(os.path.dirname(os.path.abspath(__file__)), "queries")):
    os.makedirs(out_dir, exist_ok=True)
    for name, (n, edges) in SHAPES.items():
        with open(os.path.join(out_dir, f"{name}.g"), "w") as f:
            f.write(tve(n, edges))


if __name__ == "__main__":
    main()
