"""Writes reference_binding_names.json: the names the reference's pybind11
module _mumaxpluscpp registers (src/bindings/wrap_*.cpp): module-level
functions (m.def), classes (py::class_) and each class's methods and
properties (.def, .def_property, .def_property_readonly, .def_readonly,
.def_readwrite, .def_static), plus the submodules (m.def_submodule).

Run in a checkout that has the reference at /root/reference; the JSON is the
committed fixture (names only)."""

import glob
import json
import os
import re

SRC = "/root/reference/src/bindings"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_binding_names.json")

# template arguments may nest one level (py::class_<System, std::shared_ptr<System>>)
CLASS_RE = r'py::class_<(?:[^<>]|<[^<>]*>)*>\s*\(\s*[A-Za-z_0-9]+\s*,\s*"(?P<cls>[A-Za-z_0-9]+)"'
# a chained member: the '.' follows whitespace or ')', never an identifier
MEMBER_RE = (r'(?<![A-Za-z_0-9])\.(?:def|def_property|def_property_readonly|def_readonly|def_readwrite|def_static)'
             r'\(\s*"(?P<member>[A-Za-z_0-9]+)"')
CALL_RE = r'\b(?P<var>[A-Za-z_0-9]+)\.def\(\s*"(?P<fn>[A-Za-z_0-9]+)"'
SUBMODULE_RE = re.compile(r'\b([A-Za-z_0-9]+)\s*=\s*m\.def_submodule\(\s*"([A-Za-z_0-9]+)"')
TOKEN_RE = re.compile("|".join((CLASS_RE, MEMBER_RE, CALL_RE)))


def parse():
    """Scan each file's registrations in order: a class registration starts
    a chain, chained .def* calls add members to it, m.def adds a module-level
    function and ends the chain."""
    functions, classes, submodules = {}, {}, {}
    for path in sorted(glob.glob(os.path.join(SRC, "wrap_*.cpp"))):
        fname = os.path.basename(path)
        text = open(path).read()
        subs = {v: name for v, name in SUBMODULE_RE.findall(text)}
        current = None
        for tok in TOKEN_RE.finditer(text):
            if tok.group("cls"):
                current = classes.setdefault(tok.group("cls"), {"file": fname, "members": []})["members"]
            elif tok.group("member"):
                if current is not None and tok.group("member") not in current:
                    current.append(tok.group("member"))
            elif tok.group("var") == "m":
                functions.setdefault(tok.group("fn"), fname)
                current = None
            elif tok.group("var") in subs:
                sub = submodules.setdefault(subs[tok.group("var")], {"file": fname, "functions": []})
                if tok.group("fn") not in sub["functions"]:
                    sub["functions"].append(tok.group("fn"))
                current = None
    return {"functions": functions, "classes": classes, "submodules": submodules}


if __name__ == "__main__":
    data = parse()
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print(len(data["functions"]), "functions,", len(data["classes"]), "classes,",
          sum(len(c["members"]) for c in data["classes"].values()), "members,",
          len(data["submodules"]), "submodules")
