"""Alya input decks for the explicit SMA step.

Reads a solidz case -- `<name>.dat`, `<name>.dom.dat` (with its INCLUDEd
`.geo.dat`, `.set.dat`, `.fix.dat`) and `<name>.sld.dat` -- into the
problem dict of alya_sma_amd.problems, so that a case such as the
reference's tests/uniaxial_test runs as written.  The subset read is the
one this path uses:

  .dat       TIME_INTERVAL (t0, t1), NUMBER_OF_STEPS, TIME_COUPLING
             FROM_CRITICAL
  .dom.dat   FIELDS (FIELD = n: "node fx fy fz" nodal values)
  .geo.dat   ELEMENTS (HEX08), COORDINATES, BOUNDARIES (QUA04 face nodes
             [+ element]), MATERIALS
  .fix.dat   ON_BOUNDARIES (boundary -> code), ON_NODES (node -> code)
  .set.dat   NODE (node -> node set)
  .sld.dat   PROBLEM_DEFINITION: GRAVITY (NORM, GX, GY, GZ);
             PROPERTIES: MATERIAL, DENSITY, CONSTITUTIVE_MODEL NITINOL
             (sld_reaphy.f90: parco(1:17)) or NEOHO ANSYS mu K
             (lawst 102), RAYLEIGH_DAMPING: RAYALPHA, RAYBETA
             (sld_reaphy.f90:1046-1049), FORCE: FIELD = n (nodal forces
             ramped by t/(t1 - t0), sld_matrix.f90:249-262); NUMERICAL_TREATMENT:
             TIME_INTEGRATION (TCHAMWA_WIELGOSZ with PHITW, or central
             difference), SAFETY_FACTOR; BOUNDARY_CONDITIONS[, TRANSIENT |
             CONSTANT] with CODES, NODES rules ("c1 & c2 ... fixity values
             [, MODULE_FUNCTION= f]"), CODES, BOUNDARIES rules ("code fixbo
             v1 v2 v3 [, MODULE_FUNCTION= f]", fixbo 3 pressure or 6
             traction: sld_bouope.f90), BODYF = function (body force,
             sld_reabcs.f90:523), FOLLOWER_LOADS: ON, and FUNCTIONS
             (DISCRETE[, SMOOTH] shape tables, sld_funbou.f90)

A node's codes are the codes of the boundaries whose faces contain it
(Alya's boundary-condition extrapolation to nodes); a CODES, NODES rule
applies to the nodes whose code set equals the rule's.  Fixity digits are
per dof (x y z); a prescribed dof takes the function's table column, or
the given value without a function.

write_node_sets() writes per-step node-set histories (DISPX..Z and the
reactions FRXIX..Z on prescribed dofs) as text columns, in the spirit of
Alya's node-set output (sld_outset.f90 / sld_nodset).  Witness points
(.ker.dat WITNESS_POINTS coordinates, .sld.dat WITNESS_POINTS variables) are
written by write_witness() (alya_sma_amd.witness).
"""
import os
import re

import numpy as np


def _lines(path, seen=None):
    """Statements of an Alya file: '$' comments removed, '\\' continuations
    joined, INCLUDE files expanded (relative to the including file)."""
    seen = set() if seen is None else seen
    path = os.path.abspath(path)
    if path in seen:
        raise ValueError(f"INCLUDE cycle at {path}")
    seen.add(path)
    out, acc = [], ""
    with open(path) as f:
        for raw in f:
            line = raw.split("$", 1)[0].rstrip()
            if line.endswith("\\"):
                acc += line[:-1] + " "
                continue
            line = (acc + line).strip()
            acc = ""
            if not line:
                continue
            m = re.match(r"INCLUDE\s+(\S+)", line, re.I)
            if m:
                inc = os.path.join(os.path.dirname(path), m.group(1))
                if os.path.exists(inc):
                    out.extend(_lines(inc, seen))
                continue
            out.append(line)
    return out


def _block(lines, start, end):
    """Statements strictly between the first line starting with `start`
    and the next line starting with `end` (case-insensitive)."""
    up = [l.upper() for l in lines]
    for i, l in enumerate(up):
        # a section header, not a "KEY = value" line of the same name
        if l.startswith(start) and "=" not in l and (len(l) == len(start) or not l[len(start)].isalnum()
                                                     and l[len(start)] != "_"):
            for j in range(i + 1, len(lines)):
                if up[j].startswith(end):
                    return lines[i + 1:j]
            return lines[i + 1:]
    return []


def _num(s):
    return float(s.replace("D", "E").replace("d", "e"))


def read_case(dat_path):
    """Problem dict (alya_sma_amd.problems schema) of the case whose main
    file is `dat_path` (…/<name>.dat), plus 'run', 'node_sets' and 'name'."""
    base = dat_path[:-4] if dat_path.endswith(".dat") else dat_path
    name = os.path.basename(base)
    main = _lines(base + ".dat")
    dom = _lines(base + ".dom.dat")
    sld = _lines(base + ".sld.dat")
    ker = _lines(base + ".ker.dat") if os.path.exists(base + ".ker.dat") else []

    # ---- time (.dat)
    t0, t1, nmax = 0.0, None, None
    for l in main:
        m = re.match(r"TIME_INTERVAL\s*[:=]\s*(\S+)\s*,\s*(\S+)", l, re.I)
        if m:
            t0, t1 = _num(m.group(1)), _num(m.group(2))
        m = re.match(r"NUMBER_OF_STEPS\s*=\s*(\S+)", l, re.I)
        if m:
            nmax = int(_num(m.group(1)))

    # ---- mesh (.geo.dat)
    elems = {}
    for l in _block(dom, "ELEMENTS", "END_ELEMENTS"):
        v = l.split()
        if len(v) >= 9:
            elems[int(v[0])] = [int(x) for x in v[1:9]]
    coords = {}
    for l in _block(dom, "COORDINATES", "END_COORDINATES"):
        v = l.split()
        coords[int(v[0])] = [_num(x) for x in v[1:4]]
    bounds = {}
    for l in _block(dom, "BOUNDARIES", "END_BOUNDARIES"):
        v = l.split()
        bounds[int(v[0])] = [int(x) for x in v[1:5]]          # quad face nodes (element id follows)
    fields = {}
    fb = _block(dom, "FIELDS", "END_FIELDS")
    fid = None
    for l in fb:
        m = re.match(r"FIELD\s*=\s*(\d+)", l, re.I)
        if m:
            fid = int(m.group(1))
            fields[fid] = {}
        elif l.upper().startswith("END_FIELD"):
            fid = None
        elif fid is not None:
            v = l.split()
            fields[fid][int(v[0])] = [_num(x) for x in v[1:]]
    mats = {}
    for l in _block(dom, "MATERIALS", "END_MATERIALS"):
        v = l.split()
        mats[int(v[0])] = int(v[1])
    if not elems or not coords:
        raise ValueError(f"{name}: no ELEMENTS/COORDINATES read")
    node_ids = sorted(coords)
    nidx = {n: i for i, n in enumerate(node_ids)}
    elem_ids = sorted(elems)
    coord = np.array([coords[n] for n in node_ids], dtype=np.float64)
    lnods = np.array([[nidx[n] for n in elems[e]] for e in elem_ids], dtype=np.int32)

    # ---- boundary codes (.fix.dat) -> node code sets
    bcode = {}
    for l in _block(dom, "ON_BOUNDARIES", "END_ON_BOUNDARIES"):
        v = l.split()
        bcode[int(v[0])] = int(v[1])
    node_codes = [set() for _ in node_ids]
    for b, nodes in bounds.items():
        if b in bcode:
            for n in nodes:
                node_codes[nidx[n]].add(bcode[b])
    for l in _block(dom, "ON_NODES", "END_ON_NODES"):
        v = l.split()
        node_codes[nidx[int(v[0])]].add(int(v[1]))

    # ---- node sets (.set.dat): each node takes the last set it is listed
    # in (lnset), and a set's output node is the highest-numbered node
    # carrying it (mod_reaset.f90:502-507, lnsec(lnset(ipoin)) = ipoin over
    # ascending ipoin)
    lnset = {}
    for l in _block(dom, "NODE", "END_NODE"):
        v = l.split()
        if len(v) >= 2 and v[0].isdigit():
            lnset[int(v[0])] = int(v[1])
    node_sets = {}
    for nid in sorted(lnset):
        if lnset[nid] > 0:
            node_sets.setdefault(lnset[nid], []).append(nidx[nid])

    # ---- materials (.sld.dat)
    props, rho, law, damp = {}, {}, {}, {}
    force_field = None
    cur = 1
    for l in _block(sld, "PROPERTIES", "END_PROPERTIES"):
        # RAYLEIGH_DAMPING: RAYALPHA=a, RAYBETA=b (sld_reaphy.f90:1046-1049;
        # Alya matches keywords on their first five letters, 'RAYLE', and
        # takes absent values as 0)
        if re.match(r"RAYLE\w*", l, re.I):
            def rv(key):
                mm = re.search(key + r"\w*\s*=\s*([^\s,]+)", l, re.I)
                return _num(mm.group(1)) if mm else 0.0
            damp[cur] = (rv("RAYAL"), rv("RAYBE"))
        m = re.match(r"FORCE\s*:\s*FIELD\s*=\s*(\d+)", l, re.I)
        if m:
            force_field = int(m.group(1))
        m = re.match(r"MATERIAL\s*=\s*(\d+)", l, re.I)
        if m:
            cur = int(m.group(1))
        m = re.match(r"DENSITY\s*=\s*(\S+)", l, re.I)
        if m:
            rho[cur] = _num(m.group(1))
        m = re.match(r"CONSTITUTIVE_MODEL\s*:\s*(\w+)(.*)", l, re.I)
        if m:
            model, rest = m.group(1).upper(), m.group(2).split()
            p = np.zeros(18)
            if model.startswith("NITIN"):
                vals = [_num(x) for x in rest]
                p[:len(vals)] = vals                           # parco(1:17) in deck order
                law[cur] = 160
            elif model.startswith("NEOHO") and rest and rest[0].upper().startswith("ANSYS"):
                vals = [_num(x) for x in rest[1:]]
                p[:len(vals)] = vals                           # mu, K
                law[cur] = 102
            else:
                raise ValueError(f"{name}: constitutive model {model} {' '.join(rest[:1])} not on this path")
            props[cur] = p
    mat_ids = sorted(props)
    midx = {m: i for i, m in enumerate(mat_ids)}
    damping = None
    if any(a > 0 or b > 0 for a, b in damp.values()):
        damping = np.array([damp.get(m, (0.0, 0.0)) for m in mat_ids], np.float64)
    mat_of_elem = np.array([midx.get(mats.get(e, mat_ids[0]), 0) for e in elem_ids], dtype=np.int32)

    # ---- integrator (.sld.dat NUMERICAL_TREATMENT)
    scheme, phi, safety = 0, 1.033, 1.0
    for l in _block(sld, "NUMERICAL_TREATMENT", "END_NUMERICAL_TREATMENT"):
        m = re.match(r"TIME_INTEGRATION\s*:\s*(\w+)(.*)", l, re.I)
        if m:
            if m.group(1).upper().startswith("TCHAM"):
                scheme = 1
                mm = re.search(r"PHITW\s*=\s*(\S+)", m.group(2), re.I)
                if mm:
                    phi = _num(mm.group(1))
            else:
                scheme = 0
        m = re.match(r"SAFETY_FACTOR\s*=\s*(\S+)", l, re.I)
        if m:
            safety = _num(m.group(1))

    # ---- functions
    funcs = {}
    fblock = _block(sld, "FUNCTIONS", "END_FUNCTIONS")
    i = 0
    while i < len(fblock):
        m = re.match(r"FUNCTION_NUMBER\s*:\s*(\d+)", fblock[i], re.I)
        if m:
            fid, smooth, rows = int(m.group(1)), 0, []
            j = i + 1
            while j < len(fblock) and not fblock[j].upper().startswith("END_CONDITION"):
                t = fblock[j].upper()
                if t.startswith("TIME_SHAPE"):
                    if "DISCRETE" not in t:
                        raise ValueError(f"{name}: function {fid}: only DISCRETE shapes are on this path")
                    smooth = 1 if "SMOOTH" in t else 0
                if t.startswith("SHAPE_DEFINITION"):
                    npts = int(_num(fblock[j + 1].split()[0]))
                    for k in range(npts):
                        v = [_num(x) for x in fblock[j + 2 + k].split()]
                        rows.append(v[1:4] + [v[0]])          # (v_x, v_y, v_z, t)
                    j += 1 + npts
                j += 1
            funcs[fid] = (np.array(rows), smooth)
            i = j
        i += 1
    fids = sorted(funcs)
    fidx = {f: k for k, f in enumerate(fids)}
    maxpts = max([len(funcs[f][0]) for f in fids], default=1)
    ftab = np.zeros((max(1, len(fids)), maxpts, 4))
    fnpts = np.ones(max(1, len(fids)), np.int32)
    fsmooth = np.zeros(max(1, len(fids)), np.int32)
    for f in fids:
        rows, sm = funcs[f]
        ftab[fidx[f], :len(rows)] = rows
        fnpts[fidx[f]] = len(rows)
        fsmooth[fidx[f]] = sm

    # ---- nodal forces (FORCE: FIELD): f_ext = field * cutim / (t1 - t0),
    # a linear function through (t0, t0/(t1-t0)) and (t1, t1/(t1-t0)), whose
    # continuation past t1 (plast cutim / tlast) is the same line
    fext0, fext_func = None, -1
    if force_field is not None:
        if force_field not in fields:
            raise ValueError(f"{name}: FORCE: FIELD={force_field} has no FIELDS entry")
        if t1 is None:
            raise ValueError(f"{name}: TIME_INTERVAL missing")
        fext0 = np.zeros((len(node_ids), 3))
        for n, vals in fields[force_field].items():
            fext0[nidx[n], :len(vals[:3])] = vals[:3]
        span = t1 - t0
        ramp = np.zeros((1, max(2, ftab.shape[1]), 4))
        ramp[0, 0] = [t0 / span, 0.0, 0.0, t0]
        ramp[0, 1] = [t1 / span, 0.0, 0.0, t1]
        if ftab.shape[1] < 2:
            ftab = np.concatenate([ftab, np.zeros((ftab.shape[0], 2 - ftab.shape[1], 4))], axis=1)
        if fids:
            ftab = np.concatenate([ftab, ramp], axis=0)
            fnpts = np.append(fnpts, np.int32(2))
            fsmooth = np.append(fsmooth, np.int32(0))
        else:
            ftab, fnpts, fsmooth = ramp, np.array([2], np.int32), np.array([0], np.int32)
        fext_func = ftab.shape[0] - 1

    # ---- boundary-condition header: transient unless CONSTANT
    transient = True
    for l in sld:
        u = l.upper().replace(" ", "")
        if u.startswith("BOUNDARY_CONDITIONS"):
            transient = "CONST" not in u
            break
    follower = any(re.match(r"FOLLOWER_LOADS\s*[:=]?\s*ON", l, re.I) for l in sld)

    # ---- CODES, BOUNDARIES rules -> Neumann face loads
    face_rules = {}
    for l in _block(sld, "CODES, BOUNDARIES", "END_CODES") or _block(sld, "CODES,BOUNDARIES", "END_CODES"):
        head = l.split(",")[0].split()
        if len(head) < 2:
            continue
        code, fixbo = int(head[0]), int(head[1])
        if fixbo not in (3, 6):
            raise ValueError(f"{name}: boundary condition {fixbo} on code {code} is not on this path "
                             "(3 pressure and 6 traction are)")
        vals = [_num(x) for x in head[2:5]] + [0.0] * (5 - len(head))
        mf = re.search(r"MODULE_FUNCTION\s*=\s*(\d+)", l, re.I)
        face_rules[code] = (fixbo, vals[:3], fidx[int(mf.group(1))] if mf else -1)
    faces = None
    if face_rules:
        sel = [b for b in sorted(bounds) if bcode.get(b) in face_rules]
        lnodb = np.array([[nidx[n] for n in bounds[b]] for b in sel], dtype=np.int32).reshape(-1, 4)
        rules = [face_rules[bcode[b]] for b in sel]
        faces = dict(lnodb=lnodb, code=np.array([r[0] for r in rules], np.int32),
                     value=np.array([r[1] for r in rules], np.float64).reshape(-1, 3),
                     func=np.array([r[2] for r in rules], np.int32),
                     transient=transient, follower=follower)

    # ---- CODES, NODES rules -> fixity and load functions
    fix = np.zeros(len(node_ids), dtype=np.uint8)
    func_of_node = np.full(len(node_ids), -1, dtype=np.int32)
    for l in _block(sld, "CODES, NODES", "END_CODES") or _block(sld, "CODES,NODES", "END_CODES"):
        m = re.match(r"([\d\s&]+?)\s+([01]{3})\s+(.*)$", l)
        if not m:
            continue
        codes = set(int(c) for c in m.group(1).split("&"))
        digits = m.group(2)
        mf = re.search(r"MODULE_FUNCTION\s*=\s*(\d+)", m.group(3), re.I)
        vals = [_num(x) for x in re.split(r"[\s,]+", m.group(3).split(",")[0].strip())[:3]]
        if any(abs(v) > 0 for v in vals) and not mf:
            raise ValueError(f"{name}: constant nonzero prescribed values are not on this path")
        mask = sum(1 << d for d in range(3) if digits[d] == "1")
        for n, cs in enumerate(node_codes):
            if cs == codes:
                fix[n] = mask
                func_of_node[n] = fidx[int(mf.group(1))] if (mf and mask) else -1

    if t1 is None:
        raise ValueError(f"{name}: TIME_INTERVAL missing")
    prob = dict(coord=coord, lnods=lnods, props=np.stack([props[m] for m in mat_ids]),
                rho=np.array([rho.get(m, 1.0) for m in mat_ids]), mat_of_elem=mat_of_elem,
                law_of_mat=np.array([law[m] for m in mat_ids], np.int32), fixmask=fix,
                func_of_node=func_of_node, func_table=ftab, func_npts=fnpts, func_smooth=fsmooth,
                fext0=fext0, fext_func=fext_func)
    if faces is not None:
        prob["faces"] = faces
    if damping is not None:
        prob["damping"] = damping
    # ---- body force (BOUNDARY_CONDITIONS: BODYF = function, sld_reabcs.f90:
    # 523): bodyf_sld follows the function (sld_funbou.f90 itask 3) when the
    # conditions are transient and a node carries a solidz function
    # (sld_updbcs.f90:78-99); otherwise it stays zero
    for l in sld:
        mb = re.match(r"BODYF\w*\s*[:=]?\s*(\d+)", l, re.I)
        if mb:
            fb = int(mb.group(1))
            if fb > 0 and transient and np.any(func_of_node >= 0):
                if fb not in fidx:
                    raise ValueError(f"{name}: BODYF = {fb} names no function")
                prob["body_func"] = fidx[fb]
            break
    # ---- gravity (.sld.dat PROBLEM_DEFINITION, sld_reaphy.f90:209-223):
    # g = NORM times the normalised (GX, GY, GZ), each 0 when absent
    for l in _block(sld, "PROBLEM_DEFINITION", "END_PROBLEM_DEFINITION"):
        if re.match(r"GRAVI\w*\s*:", l, re.I):
            val = lambda k: next((_num(mm.group(1)) for mm in
                                  [re.search(k + r"\w*\s*=\s*([^\s,]+)", l, re.I)] if mm), 0.0)
            d = np.array([val("GX"), val("GY"), val("GZ")])
            nrm = float(np.linalg.norm(d))
            if nrm > 0.0:
                d = d / nrm
            g = val("NORM") * d
            if np.any(g != 0.0):
                prob["gravity"] = g
    # ---- witness points (.ker.dat coordinates, .sld.dat variables)
    wpts = []
    for i, l in enumerate(ker):
        if re.match(r"WITNESS_POINTS\b", l, re.I) and not re.search(r"GEOME|MESH", l, re.I):
            sc = [1.0, 1.0, 1.0]
            for k, key in enumerate(("XSCAL", "YSCAL", "ZSCAL")):
                mm = re.search(key + r"\w*\s*=\s*([^\s,]+)", l, re.I)
                if mm:
                    sc[k] = _num(mm.group(1))
            for row in ker[i + 1:]:
                if row.upper().startswith("END_WITNESS"):
                    break
                v = [_num(x) for x in re.split(r"[\s,]+", row.strip())[:3]]
                wpts.append([sc[k] * v[k] for k in range(3)])
            break
    wvars = [l.split()[0].upper() for l in _block(sld, "WITNESS_POINTS", "END_WITNESS_POINTS")]
    if wpts and wvars:
        from .witness import VARIABLES
        prob["witness"] = dict(points=np.array(wpts, np.float64),
                               variables=[v for v in wvars if v in VARIABLES])
    prob["run"] = dict(t0=t0, t_end=t1, max_steps=nmax, scheme=scheme, phi=phi, safety=safety)
    prob["node_sets"] = {k: np.array(v, np.int64) for k, v in sorted(node_sets.items())}
    prob["set_node"] = {k: int(v[-1]) for k, v in sorted(node_sets.items())}
    prob["node_codes"] = node_codes
    prob["node_ids"] = np.array(node_ids)
    prob["name"] = name
    return prob


def reactions(prob, fint, mass=None):
    """FRXIX..Z = -frxid_sld on prescribed dofs, zero elsewhere: frxid is
    elrhs = elfext - elfint there (sld_elmmat.f90:228-244), so f_int, less
    the gravity load m g when the case has gravity (mass = the lumped mass)."""
    mask = np.stack([(prob["fixmask"] >> d) & 1 for d in range(3)], axis=1).astype(bool)
    g = prob.get("gravity")
    r = fint if g is None or mass is None else fint - np.asarray(mass)[:, None] * g
    return np.where(mask, r, 0.0)


def write_node_sets(path, prob, history):
    """history: list of (step, time, u [npoin,3], frx [npoin,3]) with frx =
    -frxid_sld (FRXIX..Z, sld_nodset.f90:98-103).  One row per (step, set):
    DISPX..Z and FRXIX..Z of the set's output node (sld_nodset.f90:74-103)."""
    with open(path, "w") as f:
        f.write("# node sets of " + prob["name"] + "\n")
        f.write("# columns: step time set DISPX DISPY DISPZ FRXIX FRXIY FRXIZ\n")
        for step, t, u, frx in history:
            for s, node in prob["set_node"].items():
                d = u[node]
                q = frx[node]
                f.write(f"{step:d} {t:.15e} {s:d} " + " ".join(f"{x:.15e}" for x in (*d, *q)) + "\n")


def write_witness(path, prob, rows):
    """rows: list of (step, time, values [nw, nvar]).  One line per (step,
    point): step, time, point number (1-based), then the variables in the
    order of the deck's WITNESS_POINTS list."""
    w = prob["witness"]
    with open(path, "w") as f:
        f.write("# witness points of " + prob["name"] + "\n")
        f.write("# columns: step time point " + " ".join(w["variables"]) + "\n")
        for step, t, vals in rows:
            for k, v in enumerate(vals):
                f.write(f"{step:d} {t:.15e} {k + 1:d} " + " ".join(f"{x:.15e}" for x in v) + "\n")


def witness_fields(prob, dev):
    """Nodal fields the witness variables need, read from the device."""
    from . import witness as W
    names = prob["witness"]["variables"]
    keys = {W.VARIABLES[n][0] for n in names}
    u = dev.get("u")
    f = {"u": u}
    if "v" in keys:
        f["v"] = dev.get("v")
    if "a" in keys:
        f["a"] = dev.get("a")
    if "frx" in keys:
        f["frx"] = -dev.get("reactions")
    if "sig" in keys or "vm" in keys:
        f["sig"] = W.nodal_cauchy(prob["coord"], prob["lnods"], u, dev.get("pk1"))
        f["vm"] = W.von_mises(f["sig"])
    if "eps" in keys or "lep" in keys:
        f["eps"], f["lep"] = W.nodal_strains(prob["coord"], prob["lnods"], u)
    return f


def run_case(dat_path, out_dir=".", every=None, mode=0):
    """Run a deck on the GPU to its TIME_INTERVAL end with dt = safety x the
    critical step (TIME_COUPLING FROM_CRITICAL), writing
    <out_dir>/<name>-nodes.sld.set every `every` steps (default: 200 rows),
    and <out_dir>/<name>-witness.sld.wit when the deck has witness points."""
    from .solid import Solidz
    from . import witness as W
    prob = read_case(dat_path)
    r = prob["run"]
    wit = prob.get("witness")
    if wit is not None:
        welem, wxi = W.locate(prob["coord"], prob["lnods"], wit["points"])
    with Solidz(prob) as dev:
        dev.set_mode(mode)
        if wit is not None and any(v.startswith(("SIG", "SEQ")) for v in wit["variables"]):
            dev.set_store(True)
        dt = r["safety"] * dev.critical_dt()
        nsteps = int(np.ceil((r["t_end"] - r["t0"]) / dt))
        if r["max_steps"]:
            nsteps = min(nsteps, r["max_steps"])
        every = every or max(1, nsteps // 200)
        hist, wrows, done = [], [], 0
        while done < nsteps:
            n = min(every, nsteps - done)
            dev.step(dt, nsteps=n, scheme=r["scheme"], phi=r["phi"])
            done += n
            t = float(dev.get("time")[0])
            # FRXIX = -frxid: the device's reactions hold frxid (f_ext - f_int
            # - the damping force on prescribed dofs, gravity and body force
            # included)
            hist.append((done, t, dev.get("u"), -dev.get("reactions")))
            if wit is not None:
                wrows.append((done, t, W.values(prob["lnods"], welem, wxi, witness_fields(prob, dev),
                                                wit["variables"])))
    path = os.path.join(out_dir, prob["name"] + "-nodes.sld.set")
    write_node_sets(path, prob, hist)
    if wit is not None:
        write_witness(os.path.join(out_dir, prob["name"] + "-witness.sld.wit"), prob, wrows)
    return path, prob, dt, nsteps


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="run an Alya solidz SMA deck on the GPU")
    ap.add_argument("dat", help="path to <name>.dat")
    ap.add_argument("--out", default=".")
    ap.add_argument("--every", type=int, default=None)
    a = ap.parse_args()
    p, prob, dt, n = run_case(a.dat, a.out, a.every)
    print(f"{prob['name']}: {n} steps of {dt:.4e} s -> {p}")
