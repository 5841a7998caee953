"""Wait-count check of gfx950 assembly (test infrastructure).

For every kernel in an assembly file written by `hipcc --cuda-device-only -S`,
a forward dataflow pass over the kernel's basic blocks tracks, for each VGPR
or AGPR that an in-flight load writes, how many loads on the same counter were
issued after it (its age):

  vmcnt    global / buffer / flat / scratch loads, and the stores that share
           the counter on gfx9-family chips;
  lgkmcnt  LDS reads and writes, scalar memory loads.

`s_waitcnt vmcnt(N)` retires every register of age >= N (loads on one counter
return in order); where paths meet, the smaller age wins (the later
retirement).  A finding is an instruction that reads such a register, or
overwrites it without being a load, while the load may still be in flight on
some path into it: the value read would be stale or the load would clobber
the result.  Loads into LDS (`global_load_lds_*`, `... lds`) name an address
in operand 0, not a destination.

Used by tests/test_asm_waits.py; the full check over the library's kernels
runs when INQX_ASM_CHECK=1.
"""
import re

CAP = 64
_VLOAD = ('global_load', 'buffer_load', 'flat_load', 'scratch_load')
_VSTORE = ('global_store', 'buffer_store', 'flat_store', 'scratch_store', 'global_atomic', 'buffer_atomic')
_LDS_READ = ('ds_read', 'ds_load', 'ds_swizzle', 'ds_bpermute', 'ds_permute')
_LDS_WRITE = ('ds_write', 'ds_store', 'ds_add')


def _regs(tok):
    out = []
    for m in re.finditer(r'\b([va])\[(\d+):(\d+)\]|\b([va])(\d+)\b', tok):
        if m.group(1):
            out += [f"{m.group(1)}{k}" for k in range(int(m.group(2)), int(m.group(3)) + 1)]
        else:
            out.append(f"{m.group(4)}{m.group(5)}")
    return out


def _parse(text):
    ins, lab = [], {}
    for ln, raw in enumerate(text.split('\n'), 1):
        l = raw.split(';')[0].strip()
        if not l:
            continue
        if l.endswith(':'):
            lab[l[:-1]] = len(ins)
            continue
        if l.startswith('.'):
            continue
        op, _, rest = l.partition(' ')
        ins.append((ln, op, [o.strip() for o in rest.split(',')] if rest else [], l))
    return ins, lab


def _blocks(ins, lab):
    ends = ('s_branch', 's_endpgm', 's_setpc_b64')
    lead = {0} | set(lab.values())
    for k, (_, op, _, _) in enumerate(ins):
        if op.startswith('s_cbranch') or op in ends:
            lead.add(k + 1)
    lead = sorted(x for x in lead if x < len(ins))
    blocks = list(zip(lead, lead[1:] + [len(ins)]))
    at = {a: i for i, (a, _) in enumerate(blocks)}
    succ = []
    for a, b in blocks:
        _, op, ops, _ = ins[b - 1]
        s = []
        if op == 's_branch':
            s = [at[lab[ops[0]]]]
        elif op.startswith('s_cbranch'):
            s = [at[lab[ops[0]]]] + ([at[b]] if b < len(ins) else [])
        elif op not in ends and b < len(ins):
            s = [at[b]]
        succ.append(s)
    return blocks, succ


def _step(state, inst, found):
    ln, op, ops, text = inst
    vm, lg = state
    if op == 's_waitcnt':
        arg = ' '.join(ops)
        m = re.search(r'vmcnt\((\d+)\)', arg)
        if m:
            vm = {r: a for r, a in vm.items() if a < int(m.group(1))}
        m = re.search(r'lgkmcnt\((\d+)\)', arg)
        if m:
            lg = {r: a for r, a in lg.items() if a < int(m.group(1))}
        return vm, lg
    vload = op.startswith(_VLOAD)
    to_lds = vload and ('_lds' in op or text.endswith(' lds'))
    vstore = op.startswith(_VSTORE)
    lds_r, lds_w = op.startswith(_LDS_READ), op.startswith(_LDS_WRITE)
    smem = op.startswith(('s_load', 's_buffer_load'))
    scalar = op.startswith('s_')
    no_dst = vstore or lds_w or to_lds or scalar
    dst = [] if no_dst or not ops else _regs(ops[0])
    srcs = []
    for o in (ops if no_dst else ops[1:]):
        srcs += _regs(o)
    if 'permlane' in op and 'swap' in op:
        srcs += _regs(ops[0])
    if op.startswith('v_') and ('fmac' in op or 'mac_' in op):
        srcs += dst
    if found is not None:
        for r in srcs:
            if r in vm or r in lg:
                found.append(f"line {ln}: reads {r} in flight: {text}")
        for r in dst:
            if (r in vm or r in lg) and not (vload or lds_r):
                found.append(f"line {ln}: overwrites {r} in flight: {text}")
    if vload or vstore:
        vm = {r: min(a + 1, CAP) for r, a in vm.items()}
        if vload:
            vm.update({r: 0 for r in dst})
    if lds_r or lds_w or smem:
        lg = {r: min(a + 1, CAP) for r, a in lg.items()}
        if lds_r:
            lg.update({r: 0 for r in dst})
    return vm, lg


def _merge(x, y):
    if x is None:
        return y
    out = []
    for a, b in zip(x, y):
        d = dict(a)
        for r, v in b.items():
            d[r] = min(d.get(r, CAP), v)
        out.append(d)
    return tuple(out)


def check_kernel(text):
    """findings (strings) for the assembly of one kernel"""
    ins, lab = _parse(text)
    if not ins:
        return []
    blocks, succ = _blocks(ins, lab)
    state_in = [None] * len(blocks)
    state_in[0] = ({}, {})
    work = [0]
    while work:
        i = work.pop()
        st = state_in[i]
        for k in range(*blocks[i]):
            st = _step(st, ins[k], None)
        for s in succ[i]:
            nw = _merge(state_in[s], st)
            if nw != state_in[s]:
                state_in[s] = nw
                work.append(s)
    found = []
    for i, (a, b) in enumerate(blocks):
        st = state_in[i]
        if st is None:
            continue
        for k in range(a, b):
            st = _step(st, ins[k], found)
    return found


def check_file(asm):
    """{kernel name: findings} for every kernel of an assembly file with findings, and the kernel count"""
    names = re.findall(r'^(_Z\S+):\s*; @', asm, re.M)
    bad = {}
    for n in names:
        i = asm.index(n + ':')
        j = asm.index('.Lfunc_end', i)
        f = check_kernel(asm[i:j])
        if f:
            bad[n] = f
    return bad, len(names)
