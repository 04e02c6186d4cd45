"""Generate the swizzle member lists included inside the glsl.h vector unions.

GLSL allows any read swizzle of up to four components from the sets xyzw and
rgba (stpq is not generated: none of the scene corpus uses it).  Each member is
a proxy type sharing the parent's float storage; see glsl.h for the proxies.
Run:  python3 snelly_amd/csrc/gen_swizzle.py
"""
import itertools
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SETS = ["xyzw", "rgba"]
RESULT = {2: "vec2", 3: "vec3", 4: "vec4"}
PROXY = {2: "sw2", 3: "sw3", 4: "sw4"}


def members(n):
    out = []
    for names in SETS:
        names = names[:n]
        for k in (2, 3, 4):
            for combo in itertools.product(range(n), repeat=k):
                name = "".join(names[i] for i in combo)
                idx = ",".join(str(i) for i in combo)
                out.append(f"{PROXY[k]}<{n},{RESULT[k]},{idx}> {name};")
    return out


for n in (2, 3, 4):
    path = os.path.join(HERE, f"glsl_swz{n}.h")
    with open(path, "w") as f:
        f.write(f"// generated by gen_swizzle.py: swizzle members of a {n}-component vector\n")
        for m in members(n):
            f.write(m + "\n")
    print(path, len(members(n)))
