"""Spin-integration compiler: spin-orbital contraction programs -> block IR.

The coupled-cluster equations are written once, in spin-orbital form, as a
sequence of binary contractions over antisymmetric tensors (see
``ccsd_program.py``).  This module turns each statement into contractions
over the alpha/beta spin blocks that are actually stored -- the same blocks
ccq keeps in its t_vec (T1A/T1B, T2A/T2B/T2C, T3A..T3D;
src/utils/cc_utils.f90:479-559) -- by

1. enumerating spin assignments of every index label, dropping those that
   violate spin conservation in any factor (#beta(bra) == #beta(ket));
2. mapping every factor to its canonical stored block (indices of each
   antisymmetric group sorted: occupied before virtual, alpha before beta)
   with the permutation parity as a sign;
3. merging block contractions that are identical after renaming summed
   labels (e.g. the (e alpha, f beta) and (e beta, f alpha) halves of a
   1/2 sum_ef ladder become one contraction with weight 1);
4. for residual targets, folding the images of a term under the block's
   own antisymmetry group into one representative (the update kernel then
   applies the group antisymmetriser, R = A_G(U));
5. keeping only the block contractions a section's targets depend on.

The result is a list of ``BlockOp``s per section; ``emit.py`` writes them as
the text IR the HIP executor runs, and ``interp.py`` evaluates them with
numpy for CPU checks.
"""

from __future__ import annotations

import dataclasses
import itertools
from fractions import Fraction

OCC = set("ijklmnop")
VIR = set("abcdefghqrstuvwxyz")

ALPHA, BETA = 0, 1


def space_of(label: str) -> str:
    if label in OCC:
        return 'o'
    if label in VIR:
        return 'v'
    raise ValueError(f"label {label!r} has no orbital space")


def parity(perm) -> int:
    perm = list(perm)
    sign = 1
    seen = [False] * len(perm)
    for i in range(len(perm)):
        if seen[i]:
            continue
        j, cyc = i, 0
        while not seen[j]:
            seen[j] = True
            j = perm[j]
            cyc += 1
        if cyc % 2 == 0:
            sign = -sign
    return sign


@dataclasses.dataclass
class TensorDecl:
    name: str
    rank: int
    nbra: int                 # indices [0, nbra) are bra, the rest ket
    groups: tuple = ()        # antisymmetric position groups
    kind: str = 'inter'       # base | amp | inter | resid
    spin_sort_only: bool = False   # base integrals: canonical order by spin only

    def group_of(self, pos):
        for g in self.groups:
            if pos in g:
                return g
        return None


@dataclasses.dataclass(frozen=True)
class Ref:
    name: str
    labels: str


@dataclasses.dataclass
class Stmt:
    out: Ref
    coef: Fraction
    ins: tuple                # one or two Refs


@dataclasses.dataclass(frozen=True)
class BlockKey:
    """A stored block: tensor name, per-index spaces and spins (canonical)."""
    name: str
    spaces: str
    spins: tuple

    def tag(self) -> str:
        return f"{self.name}.{self.spaces}.{''.join('ab'[s] for s in self.spins)}"


@dataclasses.dataclass
class BlockOp:
    out: BlockKey
    out_labels: str
    coef: Fraction
    ins: tuple                # tuple of (BlockKey, labels)
    assign: bool = False      # first write to the output block in its section

    def key(self):
        return (self.out, self.out_labels, self.ins)


class Program:
    def __init__(self):
        self.decls: dict[str, TensorDecl] = {}

    def declare(self, name, rank, nbra, groups=(), kind='inter', spin_sort_only=False):
        self.decls[name] = TensorDecl(name, rank, nbra, tuple(tuple(g) for g in groups),
                                      kind, spin_sort_only)
        return name

    # ------------------------------------------------------------------
    # statement parsing: "R2[ijab] += 1/2 V[abef] * tau[ijef]"

    def stmt(self, text: str) -> Stmt:
        lhs, rhs = text.split("+=") if "+=" in text else text.split("-=")
        neg = "-=" in text
        out = self._ref(lhs.strip())
        rhs = rhs.strip()
        parts = [p.strip() for p in rhs.split("*")]
        coef = Fraction(1)
        ins = []
        for p in parts:
            if "[" in p:
                ins.append(self._ref(p))
            else:
                coef *= Fraction(p)
        if neg:
            coef = -coef
        for r in (out, *ins):
            d = self.decls[r.name]
            if len(r.labels) != d.rank:
                raise ValueError(f"rank mismatch in {text!r} for {r.name}")
        allin = "".join(r.labels for r in ins)
        for lab in out.labels:
            if lab not in allin:
                raise ValueError(f"output label {lab} not produced in {text!r}")
        return Stmt(out, coef, tuple(ins))

    @staticmethod
    def _ref(s: str) -> Ref:
        name, rest = s.split("[")
        return Ref(name.strip(), rest.rstrip("]").strip())


# ----------------------------------------------------------------------
# spin integration


def _canonical(decl: TensorDecl, labels: str, spins: dict):
    """Return (sign, BlockKey, permuted labels) for a factor with given spins."""
    order = list(range(decl.rank))
    sign = 1
    for g in decl.groups:
        pos = list(g)
        if decl.spin_sort_only:
            key = lambda p: spins[labels[p]]
        else:
            key = lambda p: (space_of(labels[p]), spins[labels[p]])
        srt = sorted(pos, key=key)
        perm = [pos.index(p) for p in srt]
        sign *= parity(perm)
        for dst, src in zip(pos, srt):
            order[dst] = src
    new_labels = "".join(labels[p] for p in order)
    key = BlockKey(decl.name, "".join(space_of(c) for c in new_labels),
                   tuple(spins[c] for c in new_labels))
    return sign, key, new_labels


def _conserves(decl: TensorDecl, labels: str, spins: dict) -> bool:
    bra = sum(spins[c] for c in labels[:decl.nbra])
    ket = sum(spins[c] for c in labels[decl.nbra:])
    return bra == ket


def _out_ok(decl: TensorDecl, labels: str, spins: dict) -> bool:
    """Whether a spin assignment of an output is the one written for its block.

    Inside each antisymmetric group, labels of equal space must carry
    non-decreasing spins; the swapped assignment describes the same stored
    block and is skipped.  Labels of different spaces are reordered (with
    sign) onto the canonical occupied-first order.
    """
    for g in decl.groups:
        for x in range(len(g)):
            for y in range(x + 1, len(g)):
                p, q = g[x], g[y]
                same = decl.spin_sort_only or space_of(labels[p]) == space_of(labels[q])
                if same and spins[labels[p]] > spins[labels[q]]:
                    return False
    return True


def _rename_dummies(out_labels: str, ins):
    """Rename summed labels in order of first appearance, keeping their space."""
    ext = set(out_labels)
    pool = {'o': iter("ijklmnop"), 'v': iter("abcdefghqrstuvwxyz")}
    used = set(ext)
    mapping = {}
    for _, labs in ins:
        for c in labs:
            if c in ext or c in mapping:
                continue
            sp = space_of(c)
            nxt = next(pool[sp])
            while nxt in used:
                nxt = next(pool[sp])
            mapping[c] = nxt
            used.add(nxt)
    new_ins = tuple((k, "".join(mapping.get(c, c) for c in labs)) for k, labs in ins)
    return new_ins


def _normalize_out(out_labels: str, ins):
    """Rename labels so that the output labels read in a fixed canonical order.

    Output labels are renamed to the first free letters of their space in
    position order; this makes ops comparable regardless of the letters the
    statement used.
    """
    pool = {'o': iter("ijklmnop"), 'v': iter("abcdefghqrstuvwxyz")}
    mapping = {}
    for c in out_labels:
        mapping[c] = next(pool[space_of(c)])
    # summed labels must not collide with the renamed externals
    tmp = {}
    taken = set(mapping.values())
    allin = "".join(l for _, l in ins)
    spare = {'o': [x for x in "ponmlkji" if x not in taken and x not in mapping],
             'v': [x for x in "zyxwvutsrqhgfedcba" if x not in taken and x not in mapping]}
    for c in allin:
        if c in mapping or c in tmp:
            continue
        tmp[c] = spare[space_of(c)].pop(0)
    full = {**mapping, **tmp}
    new_out = "".join(full[c] for c in out_labels)
    new_ins = tuple((k, "".join(full[c] for c in labs)) for k, labs in ins)
    return new_out, _rename_dummies(new_out, new_ins)


def integrate(prog: Program, st: Stmt):
    """Spin-integrate one statement; return merged BlockOps (canonical outputs)."""
    od = prog.decls[st.out.name]
    labels = sorted(set(st.out.labels + "".join(r.labels for r in st.ins)))
    ops = {}
    for bits in itertools.product((ALPHA, BETA), repeat=len(labels)):
        spins = dict(zip(labels, bits))
        if not _conserves(od, st.out.labels, spins):
            continue
        if not _out_ok(od, st.out.labels, spins):
            continue
        ok = True
        sign = 1
        ins = []
        for r in st.ins:
            d = prog.decls[r.name]
            if not _conserves(d, r.labels, spins):
                ok = False
                break
            s, key, lab = _canonical(d, r.labels, spins)
            sign *= s
            ins.append((key, lab))
        if not ok:
            continue
        osign, okey, olab = _canonical(od, st.out.labels, spins)
        nout, nins = _normalize_out(olab, tuple(ins))
        op = BlockOp(okey, nout, st.coef * sign * osign, nins)
        k = op.key()
        if k in ops:
            ops[k].coef += op.coef
        else:
            ops[k] = op
    return [op for op in ops.values() if op.coef != 0]


def group_images(decl: TensorDecl, key: BlockKey):
    """Permutations of output positions that preserve the block (same space and
    spin inside one antisymmetric group), with parity."""
    classes = []
    for g in decl.groups:
        buckets = {}
        for p in g:
            buckets.setdefault((key.spaces[p], key.spins[p]), []).append(p)
        classes += [b for b in buckets.values() if len(b) > 1]
    if not classes:
        return [(tuple(range(decl.rank)), 1)]
    out = []
    for combo in itertools.product(*[list(itertools.permutations(c)) for c in classes]):
        perm = list(range(decl.rank))
        sign = 1
        for cls, img in zip(classes, combo):
            for src, dst in zip(cls, img):
                perm[src] = dst
            sign *= parity([cls.index(x) for x in img])
        out.append((tuple(perm), sign))
    return out


def fold_images(decl: TensorDecl, ops):
    """Merge ops related by the target block's antisymmetry group (R = A_G U)."""
    remaining = {op.key(): op for op in ops}
    result = []
    for k0 in list(remaining):
        if k0 not in remaining:
            continue
        op = remaining.pop(k0)
        for perm, sign in group_images(decl, op.out):
            if perm == tuple(range(decl.rank)):
                continue
            # image: out position p receives the label formerly at perm[p]
            relabel = {op.out_labels[perm[p]]: op.out_labels[p] for p in range(decl.rank)}
            ins = tuple((k, "".join(relabel.get(c, c) for c in labs)) for k, labs in op.ins)
            nout, nins = _normalize_out(op.out_labels, _rename_dummies(op.out_labels, ins))
            k = (op.out, nout, nins)
            if k in remaining:
                other = remaining.pop(k)
                op.coef += sign * other.coef
        if op.coef != 0:
            result.append(op)
    return result


@dataclasses.dataclass
class Section:
    name: str
    kind: str                  # 'compute' | 'update' | 'copy'
    ops: list = dataclasses.field(default_factory=list)
    target: BlockKey | None = None
    cond: str = 'always'       # always | rhf | uhf
    extra: dict = dataclasses.field(default_factory=dict)


def compile_section(prog: Program, name: str, stmts, targets, fold_target=True,
                    cond='always', live_kinds=('inter',), merge_shared=0):
    """Spin-integrate ``stmts`` and keep what ``targets`` (BlockKeys) need.

    ``live_kinds`` are the tensor kinds this section computes itself; reads of
    other kinds (amplitudes, bare integrals, tensors of earlier sections) are
    inputs.
    """
    allops = []
    for st in stmts:
        allops += integrate(prog, st)
    needed = set(targets)
    # liveness to a fixed point (a block may be read by an op listed before
    # the ops that write it)
    while True:
        grew = False
        for op in allops:
            if op.out in needed:
                for k, _ in op.ins:
                    if prog.decls[k.name].kind in live_kinds and k not in needed:
                        needed.add(k)
                        grew = True
        if not grew:
            break
    keep = [op for op in allops if op.out in needed]
    # An intermediate block that no op writes is identically zero: every spin
    # assignment of its defining contraction violates conservation on some
    # factor (e.g. X[m a c i k l] = V[m a e i] T2[k l c e] with k, l alpha and
    # c beta).  Ops reading it contribute nothing; drop them, to a fixed point.
    while True:
        written = {op.out for op in keep}
        alive = [op for op in keep
                 if all(prog.decls[k.name].kind not in live_kinds or k in written
                        for k, _ in op.ins)]
        if len(alive) == len(keep):
            break
        keep = alive
    # blocks that only fed dropped ops are no longer needed
    needed = set(targets)
    while True:
        grew = False
        for op in keep:
            if op.out in needed:
                for k, _ in op.ins:
                    if prog.decls[k.name].kind in live_kinds and k not in needed:
                        needed.add(k)
                        grew = True
        if not grew:
            break
    keep = [op for op in keep if op.out in needed]
    if not any(op.out in set(targets) for op in keep):
        raise RuntimeError(f"section {name}: no op writes its targets")
    # fold residual images per target block
    out = []
    by_out = {}
    for op in keep:
        by_out.setdefault(op.out, []).append(op)
    seen = set()
    for op in keep:
        if op.out in seen:
            continue
        seen.add(op.out)
        group = by_out[op.out]
        d = prog.decls[op.out.name]
        if fold_target and d.kind == 'resid':
            group = fold_images(d, group)
            if merge_shared and d.rank >= merge_shared:
                group = merge_shared_operand(prog, d, group, name)
        else:
            group = _merge(group)
        out.append((op.out, group))
    # restore a dependency-respecting order: blocks in order of first write
    ordered = []
    for _, group in out:
        ordered += group
    ordered = _schedule(ordered)
    first = set()
    for op in ordered:
        if op.out not in first:
            op.assign = True
            first.add(op.out)
    return Section(name=name, kind='compute', ops=ordered, cond=cond,
                   extra={'targets': list(targets)})


# sizes used to bound the merged operands (those of the factoriser)
_MERGE_SIZES = {'o': 10, 'v': 70}
_MERGE_MAX = 70 ** 3 * 10


def _merge_signature(op, rkey, image, x):
    """(shared operand key, its slot roles, the other operand's roles, labels)
    of ``op`` relabelled by a residual image, with input ``x`` as the shared
    operand.  Roles: ('o', p) output position p, ('s', k) k-th summed label in
    the shared operand's order.  None when ``op`` is not a plain contraction."""
    perm, sign = image
    lc = op.out_labels
    relabel = {lc[perm[p]]: lc[p] for p in range(len(lc))}
    ins = [(k, "".join(relabel.get(c, c) for c in labs)) for k, labs in op.ins]
    (kx, lx), (ky, ly) = ins[x], ins[1 - x]
    if len(set(lx)) != len(lx) or len(set(ly)) != len(ly):
        return None
    summed = [c for c in lx if c not in lc]
    if sorted(summed) != sorted(c for c in ly if c not in lc):
        return None
    if set(lx) | set(ly) != set(lc) | set(summed) or set(lx) & set(ly) & set(lc):
        return None
    role = {c: ('o', lc.index(c)) for c in lc}
    role.update({c: ('s', k) for k, c in enumerate(summed)})
    sig_x = tuple(role[c] for c in lx)
    roles_y = tuple(role[c] for c in ly)
    return (kx, sig_x), roles_y, sign, ins, x, role


def merge_shared_operand(prog, decl, ops, section):
    """Residual terms that read the same block through the same index mapping
    (up to the residual's own antisymmetry images, R = A_G U) are merged:
    their other operands are summed into one intermediate first, so the
    block and the residual accumulator are streamed once for the group.

      U += c1 X[..] Y1[..] + c2 X[..] Y2[..]   ->   N = c1 Y1 + c2 Y2;  U += X N

    The shared operand is an amplitude block (T2, T3, T4); the merged
    intermediate must be small (at most o v^3 at the factoriser's sizes)."""
    cands = []
    for op in ops:
        best = None
        if len(op.ins) == 2:
            kinds = [prog.decls[k.name].kind for k, _ in op.ins]
            xs = [i for i in (0, 1) if kinds[i] == 'amp' and kinds[1 - i] != 'amp'] or \
                 [i for i in (0, 1) if kinds[i] == 'amp']
            for x in xs:
                for image in group_images(decl, op.out):
                    sg = _merge_signature(op, op.out, image, x)
                    if sg is None:
                        continue
                    ky = op.ins[1 - x][0]
                    size = 1
                    for c in ky.spaces:
                        size *= _MERGE_SIZES[c]
                    if size > _MERGE_MAX:
                        continue
                    key = (sg[0][0].tag(), sg[0][1], tuple(sorted(sg[1])))
                    if best is None or key < best[0]:
                        best = (key, sg)
        cands.append(best)
    groups = {}
    for i, c in enumerate(cands):
        if c is not None:
            groups.setdefault(c[0], []).append(i)
    merged = set()
    out = []
    count = [0]
    for key, members in groups.items():
        if len(members) < 2:
            continue
        (kx, sig_x), roles_y, _, _, x, _ = cands[members[0]][1]
        ref = ops[members[0]]
        # the merged operand's spaces / spins per role, in the first member's order
        spaces, spins = [], []
        for r in roles_y:
            if r[0] == 'o':
                spaces.append(ref.out.spaces[r[1]])
                spins.append(ref.out.spins[r[1]])
            else:
                slot = sig_x.index(r)
                spaces.append(kx.spaces[slot])
                spins.append(kx.spins[slot])
        count[0] += 1
        mname = f"M{decl.rank // 2}{section.split(':')[-1]}_{count[0]}"
        prog.declare(mname, len(roles_y), 0, (), kind='inter')
        mkey = BlockKey(mname, "".join(spaces), tuple(spins))
        first = True
        for i in members:
            op = ops[i]
            _, ry, sign, ins, xi, role = cands[i][1]
            inv = {v: c for c, v in role.items()}
            nl = "".join(inv[r] for r in roles_y)
            out.append(BlockOp(mkey, nl, op.coef * sign, (ins[1 - xi],), assign=first))
            first = False
            merged.add(i)
        _, _, _, ins0, x0, role0 = cands[members[0]][1]
        inv0 = {v: c for c, v in role0.items()}
        nl0 = "".join(inv0[r] for r in roles_y)
        pair = [None, None]
        pair[x0] = ins0[x0]
        pair[1 - x0] = (mkey, nl0)
        out.append(BlockOp(ref.out, ref.out_labels, Fraction(1), tuple(pair)))
    return [op for i, op in enumerate(ops) if i not in merged] + out


def _merge(ops):
    m = {}
    for op in ops:
        k = op.key()
        if k in m:
            m[k].coef += op.coef
        else:
            m[k] = dataclasses.replace(op)
    return [op for op in m.values() if op.coef != 0]


def _schedule(ops):
    """Order ops so every block is complete before it is read."""
    writers = {}
    for i, op in enumerate(ops):
        writers.setdefault(op.out, []).append(i)
    done_blocks = set()
    placed = [False] * len(ops)
    order = []
    while len(order) < len(ops):
        progress = False
        for i, op in enumerate(ops):
            if placed[i]:
                continue
            deps = [k for k, _ in op.ins if k in writers and k != op.out]
            if all(k in done_blocks for k in deps):
                placed[i] = True
                order.append(op)
                progress = True
                if all(placed[j] for j in writers[op.out]):
                    done_blocks.add(op.out)
        if not progress:
            raise RuntimeError("cyclic block dependencies")
    return order


def blocks_read(sections, kinds=('inter',), prog=None):
    out = set()
    for s in sections:
        for op in s.ops:
            for k, _ in op.ins:
                if prog is None or prog.decls[k.name].kind in kinds:
                    out.add(k)
    return out
