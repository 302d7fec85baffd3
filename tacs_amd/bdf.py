"""Reader for the subset of Nastran bulk data the solid and shell path uses.

It takes the place of pyNastran's BDF reader that the reference's mesh
loader drives (tacs/pymeshloader.py, tacs/pytacs.py), for these cards:
GRID (positions in the basic frame or a CORD2R/C/S frame defined in it),
CHEXA (8 nodes), CQUAD4, SPC1, SPC, MAT1, PSOLID, PSHELL, FORCE, PLOAD4,
PLOAD2 and LOAD, in small-field (8-character), large-field
(16-character, ``*`` cards) and free-field (comma separated) form with
continuation lines, plus the case control's SUBCASE / SUBTITLE / LABEL /
SPC / LOAD entries.  Other bulk cards are kept as raw field lists in
``BDF.unsupported`` (the loader reports them) and never interpreted.
"""
import re

import numpy as np

_EXP = re.compile(r"^([+-]?[0-9.]+)([+-][0-9]+)$")


def nastran_float(s):
    """A Nastran real field: "7.+10" is 7e10, "1.-5" is 1e-5, "D" exponents
    are accepted; blank gives None."""
    s = s.strip()
    if not s:
        return None
    t = s.upper().replace("D", "E")
    try:
        return float(t)
    except ValueError:
        m = _EXP.match(t)
        if m:
            return float(f"{m.group(1)}E{m.group(2)}")
        raise ValueError(f"not a Nastran real: {s!r}")


def nastran_int(s):
    s = s.strip()
    return int(s) if s else None


class Card:
    """A bulk data card: name and its fields (strings, continuation fields
    appended in order; the continuation markers are dropped)."""

    __slots__ = ("name", "fields", "line", "comment")

    def __init__(self, name, fields, line):
        self.name, self.fields, self.line = name, fields, line
        self.comment = None

    def f(self, i):
        return self.fields[i] if i < len(self.fields) else ""

    def int(self, i, default=None):
        v = nastran_int(self.f(i))
        return default if v is None else v

    def float(self, i, default=None):
        v = nastran_float(self.f(i))
        return default if v is None else v


def _split_line(line):
    """Returns (name, data fields, is_continuation) of one bulk data line."""
    if "," in line.split("$")[0]:
        parts = [p.strip() for p in line.split("$")[0].rstrip("\n").split(",")]
        head = parts[0]
        cont = head.startswith("+") or head.startswith("*") or head == ""
        return head.rstrip("*").upper(), parts[1:9], cont
    text = line.rstrip("\n").expandtabs(8)
    head = text[:8].strip()
    cont = head.startswith("+") or head.startswith("*") or head == ""
    large = head.endswith("*") or head.startswith("*")
    if large:
        body = text[8:72].ljust(64)
        fields = [body[16 * k:16 * k + 16].strip() for k in range(4)]
    else:
        body = text[8:72].ljust(64)
        fields = [body[8 * k:8 * k + 8].strip() for k in range(8)]
    return head.rstrip("*").upper(), fields, cont


class BDF:
    """Parsed model.  Nodes and elements keep the file order (the order the
    reference's loader numbers them in); ids map to 0-based TACS numbers."""

    SUPPORTED = ("GRID", "CORD2R", "CORD2C", "CORD2S", "CHEXA", "CQUAD4", "SPC1", "SPC",
                 "MAT1", "PSOLID", "PSHELL", "FORCE", "PLOAD4", "PLOAD2", "LOAD")

    def __init__(self, path):
        self.path = path
        self.nodes = {}        # nid -> xyz in the basic frame
        self.elements = {}     # eid -> (type, pid, [nids] in Nastran order)
        self.spcs = {}         # sid -> list of (nids, components string, enforced)
        self.materials = {}    # mid -> dict
        self.properties = {}   # pid -> dict
        self.comments = {}     # pid -> component name from the preceding comment
        self.loads = {}        # sid -> list of load dicts
        self.load_combos = {}  # sid -> (scale, [(scale_i, sid_i)])
        self.subcases = []     # dicts: id, name, spc, load
        self.coords = {}       # cid -> (kind R/C/S, origin, axes as columns)
        self.unsupported = []
        self._read(path)
        self.nodes = {n: self._to_basic(cp, x) for n, (cp, x) in self.nodes.items()}
        self.node_ids = list(self.nodes)
        self.nid_to_tacs = {n: i for i, n in enumerate(self.node_ids)}
        self.elem_ids = list(self.elements)
        self.eid_to_tacs = {e: i for i, e in enumerate(self.elem_ids)}

    # -- reading ------------------------------------------------------------
    def _read(self, path):
        with open(path) as fh:
            lines = fh.readlines()
        bulk_start = None
        for k, line in enumerate(lines):
            if line.strip().upper().startswith("BEGIN BULK"):
                bulk_start = k + 1
                break
        case = lines[:bulk_start] if bulk_start is not None else []
        bulk = lines[bulk_start:] if bulk_start is not None else lines
        self._read_case_control(case)
        cards = []
        comment = None
        for k, line in enumerate(bulk):
            s = line.strip()
            if not s:
                continue
            if s.startswith("$"):
                comment = s[1:].strip()
                continue
            if s.upper().startswith("ENDDATA"):
                break
            name, fields, cont = _split_line(line)
            if cont:
                if not cards:
                    raise ValueError(f"{path}: continuation line without a card: {line!r}")
                cards[-1].fields.extend(fields)
            else:
                cards.append(Card(name, fields, k + 1 + (bulk_start or 0)))
                # the comment line right before a property card names the
                # component (its last word, as the reference's loader takes it)
                if name in ("PSHELL", "PSOLID") and comment:
                    cards[-1].comment = comment.split()[-1]
                comment = None
        for c in cards:
            # trailing blank fields carry no data
            while c.fields and c.fields[-1] == "":
                c.fields.pop()
            handler = getattr(self, "_card_" + c.name, None)
            if handler is not None:
                handler(c)
            elif not c.name.startswith("PARAM"):
                self.unsupported.append(c)

    def _read_case_control(self, lines):
        cur = None
        for line in lines:
            s = line.split("$")[0].strip()
            u = s.upper()
            if u.startswith("SUBCASE"):
                cur = {"id": int(u.split()[1]), "name": None, "label": None, "spc": None,
                       "load": None}
                self.subcases.append(cur)
            elif cur is not None and "=" in s:
                key, val = [x.strip() for x in s.split("=", 1)]
                key = key.upper()
                if key == "SUBTITLE":
                    cur["name"] = val
                elif key == "LABEL":
                    cur["label"] = val
                elif key == "SPC":
                    cur["spc"] = int(val)
                elif key == "LOAD":
                    cur["load"] = int(val)

    # -- cards ----------------------------------------------------------------
    def _card_GRID(self, c):
        if c.int(5, 0) != 0:
            raise ValueError(f"GRID {c.int(0)}: displacement frame CD={c.f(5)} not supported")
        self.nodes[c.int(0)] = (c.int(1, 0),
                                np.array([c.float(2, 0.0), c.float(3, 0.0), c.float(4, 0.0)]))

    def _coord(self, c, kind):
        """CORD2R / CORD2C / CORD2S given in the basic frame (RID 0): origin A,
        z axis towards B, x axis in the plane of C."""
        if c.int(1, 0) != 0:
            raise ValueError(f"CORD2{kind} {c.int(0)}: reference frame RID={c.f(1)} not supported")
        A = np.array([c.float(k, 0.0) for k in (2, 3, 4)])
        B = np.array([c.float(k, 0.0) for k in (5, 6, 7)])
        C = np.array([c.float(k, 0.0) for k in (8, 9, 10)])
        z = (B - A) / np.linalg.norm(B - A)
        y = np.cross(z, C - A)
        y /= np.linalg.norm(y)
        x = np.cross(y, z)
        self.coords[c.int(0)] = (kind, A, np.column_stack([x, y, z]))

    def _card_CORD2R(self, c):
        self._coord(c, "R")

    def _card_CORD2C(self, c):
        self._coord(c, "C")

    def _card_CORD2S(self, c):
        self._coord(c, "S")

    def _to_basic(self, cp, x):
        if cp == 0:
            return x
        if cp not in self.coords:
            raise ValueError(f"{self.path}: coordinate system {cp} is not defined")
        kind, A, R = self.coords[cp]
        if kind == "C":
            th = np.deg2rad(x[1])
            x = np.array([x[0] * np.cos(th), x[0] * np.sin(th), x[2]])
        elif kind == "S":
            th, ph = np.deg2rad(x[1]), np.deg2rad(x[2])
            x = x[0] * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        return A + R @ x

    def _card_CHEXA(self, c):
        nids = [c.int(k) for k in range(2, 10)]
        extra = [x for x in c.fields[10:] if x.strip()]
        if extra or any(n is None for n in nids):
            raise ValueError(f"CHEXA {c.int(0)}: only 8-node CHEXA is supported")
        self.elements[c.int(0)] = ("CHEXA", c.int(1), nids)

    def _card_CQUAD4(self, c):
        self.elements[c.int(0)] = ("CQUAD4", c.int(1), [c.int(k) for k in range(2, 6)])

    def _card_SPC1(self, c):
        comps = c.f(1).strip()
        rest = [x.strip() for x in c.fields[2:] if x.strip()]
        if len(rest) == 3 and rest[1].upper() == "THRU":
            nids = list(range(int(rest[0]), int(rest[2]) + 1))
        else:
            nids = [int(x) for x in rest]
        self.spcs.setdefault(c.int(0), []).append((nids, comps, 0.0))

    def _card_SPC(self, c):
        sid = c.int(0)
        for k in (1, 4):
            if c.f(k).strip():
                self.spcs.setdefault(sid, []).append(([c.int(k)], c.f(k + 1).strip(),
                                                      c.float(k + 2, 0.0)))

    def _card_MAT1(self, c):
        E, G, nu = c.float(1), c.float(2), c.float(3)
        if E is None and G is not None and nu is not None:
            E = 2.0 * (1.0 + nu) * G
        if nu is None and E is not None and G is not None:
            nu = E / (2.0 * G) - 1.0
        self.materials[c.int(0)] = {"E": E, "nu": nu if nu is not None else 0.0,
                                    "rho": c.float(4, 0.0), "alpha": c.float(5, 0.0),
                                    "St": c.float(8, 0.0)}

    def _card_PSOLID(self, c):
        self.properties[c.int(0)] = {"type": "PSOLID", "mid": c.int(1)}
        if c.comment:
            self.comments[c.int(0)] = c.comment

    def _card_PSHELL(self, c):
        self.properties[c.int(0)] = {"type": "PSHELL", "mid": c.int(1), "t": c.float(2),
                                     "tst": c.float(6, 0.833333)}
        if c.comment:
            self.comments[c.int(0)] = c.comment

    def _card_FORCE(self, c):
        if c.int(2, 0) != 0:
            raise ValueError(f"FORCE {c.int(0)}: coordinate system CID={c.f(2)} not supported")
        vec = np.array([c.float(4, 0.0), c.float(5, 0.0), c.float(6, 0.0)]) * c.float(3, 0.0)
        self.loads.setdefault(c.int(0), []).append({"type": "FORCE", "node": c.int(1),
                                                    "vector": vec})

    def _card_PLOAD4(self, c):
        p = [c.float(k) for k in range(2, 6)]
        p = [v for v in p if v is not None]
        eids = [c.int(1)]
        if c.f(6).strip().upper() == "THRU":
            eids = list(range(c.int(1), c.int(7) + 1))
            g1 = g34 = None
        else:
            g1, g34 = c.int(6), c.int(7)
        if any(x.strip() for x in c.fields[8:]):
            raise ValueError(f"PLOAD4 {c.int(0)}: direction fields (CID, N1-N3) not supported")
        self.loads.setdefault(c.int(0), []).append({"type": "PLOAD4", "eids": eids,
                                                    "pressures": p, "g1": g1, "g34": g34})

    def _card_PLOAD2(self, c):
        rest = [x.strip() for x in c.fields[2:] if x.strip()]
        if len(rest) == 3 and rest[1].upper() == "THRU":
            eids = list(range(int(rest[0]), int(rest[2]) + 1))
        else:
            eids = [int(x) for x in rest]
        self.loads.setdefault(c.int(0), []).append({"type": "PLOAD2", "eids": eids,
                                                    "pressure": c.float(1, 0.0)})

    def _card_LOAD(self, c):
        pairs = []
        k = 2
        while k + 1 < len(c.fields) and c.f(k).strip():
            pairs.append((c.float(k), c.int(k + 1)))
            k += 2
        self.load_combos[c.int(0)] = (c.float(1, 1.0), pairs)

    # -- queries ----------------------------------------------------------------
    def reduced_loads(self, sid, scale=1.0):
        """[(load dict, scale)] of a load set, LOAD combinations expanded."""
        if sid in self.load_combos:
            s0, pairs = self.load_combos[sid]
            out = []
            for si, li in pairs:
                out.extend(self.reduced_loads(li, scale * s0 * si))
            return out
        return [(ld, scale) for ld in self.loads.get(sid, [])]
