"""Write compiled block programs as the text IR the HIP executor loads.

Format (one record per line, '#' comments):

  ccq-ir 1 <method> <rhf|uhf>
  tensor <id> inter <rank> <dim>...          dense C-order intermediate
  tensor <id> view <base> <rank> <sp>...     slice of a full-range base array
  tensor <id> amp <block>                    amplitude block (t1a, t2ab, ...)
  tensor <id> resid <block>                  residual accumulator of a block
  anti <id> <p>,<q>...                       position pairs of equal space and
                                             spin in which the stored block is
                                             antisymmetric (not for residual
                                             accumulators, which hold one
                                             representative per image set)
  program <name>                             sections that follow belong to
                                             program <name> (default: sweep)
  section <name>
  op <out> <assign> <coef> <nin> <in> <labels> [<in> <labels>] <out-labels>
  update <block> <resid-id> <nimages> [<perm>:<sign>]... [mask <vneed> <vforbid>] [left]
  copy <dst-block> <src-block> <perm>...
  cct3 <block> <l3-id> <m3-id> <nimages> [<perm>:<sign>]...
  end

Dimension symbols are oa ob va vb (correlated occupied / virtual per spin);
view index specs are two characters, space (o|v) and spin (a|b).  Base
arrays: Vaa, Vab, Vbb (<pq||rs> over the correlated orbitals, C order
[p][q][r][s]; Vab = <p q~|r s~>), Fa, Fb (Fock, C order [p][q]).
"""

from __future__ import annotations

import os

from .interp import AMP_BLOCKS, RES_OF
from .spinint import group_images

DIMSYM = {('o', 0): 'oa', ('o', 1): 'ob', ('v', 0): 'va', ('v', 1): 'vb'}


def _base_of(key):
    if key.name == 'F':
        return 'Fa' if key.spins[0] == 0 else 'Fb'
    sp = key.spins
    if sp == (0, 0, 0, 0):
        return 'Vaa'
    if sp == (1, 1, 1, 1):
        return 'Vbb'
    if sp == (0, 1, 0, 1):
        return 'Vab'
    raise KeyError(key)


def anti_pairs(decl, key):
    """Position pairs (p < q) of one declared antisymmetric group whose indices
    have the same space and spin in this block."""
    out = []
    for g in decl.groups:
        for i, p in enumerate(g):
            for q in g[i + 1:]:
                if key.spaces[p] == key.spaces[q] and key.spins[p] == key.spins[q]:
                    out.append((min(p, q), max(p, q)))
    return sorted(out)


def emit(prog, sections, method: str, rhf: bool) -> str:
    ids = {}
    lines = [f"ccq-ir 1 {method} {'rhf' if rhf else 'uhf'}"]
    res_block = {v: k for k, v in RES_OF.items()}

    def tid(key):
        if key in ids:
            return ids[key]
        kind = prog.decls[key.name].kind
        name = key.tag()
        ids[key] = name
        rank = len(key.spaces)
        if kind == 'hbar':
            kind = 'inter'
        if kind == 'base':
            specs = " ".join(c + "ab"[s] for c, s in zip(key.spaces, key.spins))
            lines.append(f"tensor {name} view {_base_of(key)} {rank} {specs}")
        elif kind == 'amp':
            lines.append(f"tensor {name} amp {AMP_BLOCKS[(key.name, key.spins)]}")
        elif kind == 'resid':
            lines.append(f"tensor {name} resid {res_block[(key.name, key.spins)]}")
        else:
            dims = " ".join(DIMSYM[(c, s)] for c, s in zip(key.spaces, key.spins))
            lines.append(f"tensor {name} inter {rank} {dims}")
        if kind != 'resid':
            pairs = anti_pairs(prog.decls[key.name], key)
            if pairs:
                lines.append(f"anti {name} " + " ".join(f"{p},{q}" for p, q in pairs))
        return name

    body = []
    program = 'sweep'
    for sec in sections:
        sp = getattr(sec, 'program', None) or sec.extra.get('program', 'sweep')
        if sp != program:
            body.append(f"program {sp}")
            program = sp
        body.append(f"section {sec.name}")
        if sec.kind == 'compute':
            for op in sec.ops:
                out = tid(op.out)
                ins = " ".join(f"{tid(k)} {labs}" for k, labs in op.ins)
                coef = float(op.coef)
                body.append(f"op {out} {int(op.assign)} {coef!r} {len(op.ins)} {ins} {op.out_labels}")
        elif sec.kind == 'update':
            blk = sec.extra['block']
            rid = tid(sec.target)
            decl = prog.decls[sec.target.name]
            imgs = group_images(decl, sec.target)
            perms = " ".join(",".join(str(x) for x in p) + ":" + str(s) for p, s in imgs)
            mask = sec.extra.get('mask')
            tail = f" mask {mask[0]} {mask[1]}" if mask is not None else ""
            if sec.extra.get('left'):
                tail += " left"
            body.append(f"update {blk} {rid} {len(imgs)} {perms}{tail}")
        elif sec.kind == 'cct3':
            lid = tid(sec.target)
            mid = tid(sec.extra['m3'])
            decl = prog.decls[sec.target.name]
            imgs = group_images(decl, sec.target)
            perms = " ".join(",".join(str(x) for x in p) + ":" + str(s) for p, s in imgs)
            body.append(f"cct3 {sec.extra['block']} {lid} {mid} {len(imgs)} {perms}")
        elif sec.kind == 'copy':
            body.append(f"copy {sec.extra['dst']} {sec.extra['src']}")
    return "\n".join(lines + body + ["end"]) + "\n"


def write_cxx_include(path: str, text: str, symbol: str, gh: str = "") -> None:
    """Embed an IR text as a C++ raw string literal (first line records the generator hash)."""
    with open(path, "w") as f:
        f.write(f"// generated by ccq_amd/gen/emit.py (generator {gh}) -- do not edit\n")
        f.write(f'static const char {symbol}[] = R"CCQIR(\n{text})CCQIR";\n')


METHODS = ("CCSD", "CCSDT", "CCSDTQ", "CCSDt", "CCT3")
# file stem of each method's program ("CCSDt" is the active-space CCSDT path,
# which ccq runs without closed-shell shortcuts: uhf only)
STEM = {"CCSD": "ccsd", "CCSDT": "ccsdt", "CCSDTQ": "ccsdtq", "CCSDt": "ccsdt_act",
        "CCT3": "cct3"}
VARIANTS = {"CCSD": (False, True), "CCSDT": (False, True), "CCSDTQ": (False, True),
            "CCSDt": (False,), "CCT3": (False,)}


def generator_hash(method: str = "CCT3") -> str:
    """Hash of the generator sources a method's program is made from; an .inc
    made from the same sources is current.  The CC(t;3) tail program
    (cct3_program.py) feeds only the CCT3 include."""
    import hashlib
    here = os.path.dirname(os.path.abspath(__file__))
    h = hashlib.sha256()
    for name in sorted(os.listdir(here)):
        if name == "cct3_program.py" and method != "CCT3":
            continue
        if name.endswith(".py"):
            with open(os.path.join(here, name), "rb") as f:
                h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def generate_all(outdir: str, methods=METHODS, force: bool = False) -> list:
    """Write ir_<stem>_<rhf|uhf>.inc for every method.

    Generation of the quadruples program takes minutes, so an existing file
    whose recorded generator hash matches the current sources is kept.
    """
    from . import cc_program
    written = []
    for method in methods:
        gh = generator_hash(method)
        for rhf in VARIANTS[method]:
            tag = "rhf" if rhf else "uhf"
            p = os.path.join(outdir, f"ir_{STEM[method]}_{tag}.inc")
            sym = f"kIR_{STEM[method].upper()}_{tag.upper()}"
            if not force and os.path.exists(p):
                with open(p) as f:
                    first = f.readline()
                if f"generator {gh}" in first:
                    written.append(p)
                    continue
            if method == "CCT3":
                from . import cct3_program
                prog, groups = cct3_program.build()
                secs = []
                for name in ('hbar', 'left', 'cct3'):
                    for sec in groups[name]:
                        sec.extra['program'] = name
                        secs.append(sec)
            else:
                prog, secs = cc_program.build(method, rhf)
            text = emit(prog, secs, method, rhf)
            write_cxx_include(p, text, sym, gh)
            written.append(p)
    return written


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    for p in generate_all(os.path.join(here, "..", "csrc")):
        print("wrote", p)
