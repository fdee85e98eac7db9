"""Index programs: which global gradient slots an element touches.

An energy's index program mirrors the path dictionary the reference builds from the
energy's expression (``differentiator/path.pyx:103-213``): DATA attributes reached
from the energy become leaves, JOIN attributes loop over their connectivity, UNION
attributes pick the child primitive of each union instance and add one padding slot
(``kernel/Coordinate/gradientIndicesKernel.pyx:546-600``). CONSTANT attributes have no
path and produce no slot.

``encode`` turns a program into the int32 bytecode that ``ysp_get_indices`` interprets
on the device (layout documented in ``include/yasps_amd.h``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

OP_LEAF, OP_JOIN, OP_UNION, OP_CONCAT = 0, 1, 2, 3


@dataclass
class PLeaf:
    size: int
    wrt_pos: Optional[int]
    per_instance: bool = True

    def sizes(self):
        return self.size, 1


@dataclass
class PJoin:
    conn: object            # connectivity (host ``value_host`` [N, dim], device ``value``)
    dim: int
    children: List[object]

    def sizes(self):
        return (sum(c.sizes()[0] for c in self.children) * self.dim,
                sum(c.sizes()[1] for c in self.children) * self.dim)


@dataclass
class PUnion:
    union: object           # primitiveUnion (host ``children_counts``, device ``counts_device``)
    members: List[List[object]]

    def sizes(self):
        g = k = 0
        for m in self.members:
            g = max(g, sum(n.sizes()[0] for n in m))
            k = max(k, sum(n.sizes()[1] for n in m) + 1)
        return g, k


@dataclass
class PConcat:
    children: List[object] = field(default_factory=list)

    def sizes(self):
        return (sum(c.sizes()[0] for c in self.children),
                sum(c.sizes()[1] for c in self.children))


def encode(node, conns: list, unions: list) -> List[int]:
    """Bytecode of ``node``; ``conns`` / ``unions`` collect the referenced objects in
    slot order."""
    g, k = node.sizes()
    if isinstance(node, PLeaf):
        return [OP_LEAF, 1, node.size, 6, -1 if node.wrt_pos is None else int(node.wrt_pos),
                1 if node.per_instance else 0]
    if isinstance(node, PJoin):
        if node.conn not in conns:
            conns.append(node.conn)
        body = []
        for c in node.children:
            body += encode(c, conns, unions)
        head = [OP_JOIN, k, g, 0, conns.index(node.conn), int(node.dim), len(node.children)]
        head[3] = len(head) + len(body)
        return head + body
    if isinstance(node, PUnion):
        if node.union not in unions:
            unions.append(node.union)
        body = []
        for members in node.members:
            body.append(len(members))
            for m in members:
                body += encode(m, conns, unions)
        head = [OP_UNION, k, g, 0, unions.index(node.union), len(node.members)]
        head[3] = len(head) + len(body)
        return head + body
    if isinstance(node, PConcat):
        body = []
        for c in node.children:
            body += encode(c, conns, unions)
        head = [OP_CONCAT, k, g, 0, len(node.children)]
        head[3] = len(head) + len(body)
        return head + body
    raise TypeError(f"unknown program node {node!r}")


def interpret(code, conns_host, counts_host, wrt_start, num_instances):
    """Host interpretation of the bytecode (same rules as the device interpreter); used
    to check the encoding against the oracle on the CPU."""
    import numpy as np
    K = code[1]
    out_i = np.zeros((num_instances, K), dtype=np.int64)
    out_s = np.zeros((num_instances, K), dtype=np.int64)

    def emit(pc, idx, oi, os, off):
        op = code[pc]
        if op == OP_LEAF:
            size, wrt_pos, per = code[pc + 2], code[pc + 4], code[pc + 5]
            oi[off] = wrt_start[wrt_pos] + idx * (size if per else 0) + 2 if wrt_pos >= 0 else 1
            os[off] = size
        elif op == OP_JOIN:
            k, conn, dim, nch = code[pc + 1], conns_host[code[pc + 4]], code[pc + 5], code[pc + 6]
            per = k // dim
            flat = np.asarray(conn).reshape(-1)
            for i in range(dim):
                target = int(flat[idx * dim + i])
                cpc, acc = pc + 7, 0
                for _ in range(nch):
                    emit(cpc, target, oi, os, off + per * i + acc)
                    acc += code[cpc + 1]
                    cpc += code[cpc + 3]
        elif op == OP_UNION:
            k, g, counts, nuc = code[pc + 1], code[pc + 2], counts_host[code[pc + 4]], code[pc + 5]
            total, which = 0, 0
            for c in range(nuc):
                if total + counts[c] > idx:
                    which = c
                    break
                total += counts[c]
            cpc = pc + 6
            for _ in range(which):
                nm = code[cpc]
                cpc += 1
                for _ in range(nm):
                    cpc += code[cpc + 3]
            nm = code[cpc]
            cpc += 1
            shift = actual = 0
            for _ in range(nm):
                emit(cpc, idx - total, oi, os, off + shift)
                shift += code[cpc + 1]
                actual += code[cpc + 2]
                cpc += code[cpc + 3]
            os[off + k - 1] = g - actual
        elif op == OP_CONCAT:
            cpc, acc = pc + 5, 0
            for _ in range(code[pc + 4]):
                emit(cpc, idx, oi, os, off + acc)
                acc += code[cpc + 1]
                cpc += code[cpc + 3]
        else:
            raise ValueError(f"bad opcode {op}")

    for e in range(num_instances):
        emit(0, e, out_i[e], out_s[e], 0)
    return out_i.astype(np.uint32), out_s.astype(np.uint16)
