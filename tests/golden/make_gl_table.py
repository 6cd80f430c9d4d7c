"""Extract the Gauss-Lobatto (Solin) known-answer table shipped with the
reference, MONTJOIE/src/Quadrature/Table_Quadrature.txt:172-271 (points on
[-1,1], weights summing to 2, 25 digits), into tests/golden/gl_solin.json.
Run in the build container only (reads /root/reference)."""
import json
import os

SRC = "/root/reference/MONTJOIE/src/Quadrature/Table_Quadrature.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gl_solin.json")

lines = open(SRC).read().splitlines()
# locate the Lobatto block: entries "<n>" then n lines "point weight", each
# block followed by a "Lobatto  (Radau)  quadrature constants" caption
rules = {}
i = 0
while i < len(lines):
    s = lines[i].strip()
    if s.isdigit():
        n = int(s)
        pts = []
        ok = True
        for k in range(n):
            parts = lines[i + 1 + k].split()
            if len(parts) != 2:
                ok = False
                break
            pts.append([parts[0], parts[1]])
        cap = lines[i + 1 + n].strip() if ok and i + 1 + n < len(lines) else ""
        if ok and cap.startswith("Lobatto"):
            rules[str(n)] = pts
            i += n + 2
            continue
    i += 1
json.dump({"source": "MONTJOIE/src/Quadrature/Table_Quadrature.txt:172-271",
           "domain": "[-1,1], weights sum to 2", "rules": rules}, open(OUT, "w"), indent=1)
print("rules:", sorted(map(int, rules)))
