"""Restatement of the reference's edge-list parser and CSR builder.

TEST INFRASTRUCTURE ONLY.  Follows, step by step:

* ``kites/graph/graph_raw.h:73-123`` -- read line by line; skip a line whose
  first character is not a digit (``:74-75``); split on space and tab only
  (``:67``, so a CRLF line keeps ``\\r`` on its last token and ``atoll`` stops
  there); the first two tokens are ``atoll``'d ids (``:80-85``); further
  tokens go to the callback as ``argv`` (``:96-101``); vectors grow to
  ``max id + 1`` with value-initialised entries (``:90-94``); each edge is an
  in-neighbour of its destination (``:102``) and the callback's return value
  is OR-ed into the initial frontier bit of the destination's group
  (``:103-110``); undirected + incomplete adds the reverse edge with the
  callback called again on swapped ends (``:111-121``).
* ``graph_raw.h:126-151`` -- sort each in-list by source id, keeping
  duplicates (``:131-133``); directed graphs get group-level out-lists
  (``:135-139``) that are sorted and de-duplicated only when vpg != 1
  (``:142-151``).
* ``graph_raw.h:173-182`` -- pad the vertex count to a multiple of 128; the
  out-lists and the frontier are padded only for directed graphs with vpg 1.
* ``graph_csr.h:77-108`` -- fill V, R, C, E; the frontier bits and dir_R /
  dir_C come from the loop over the out-list groups (``:100-108``), which is
  empty for undirected graphs (the reference therefore starts undirected
  graphs with an empty frontier; ``frontier_as_reference`` reproduces that,
  ``frontier`` keeps the bits the callback produced).

The reference sorts with ``std::sort`` (not stable); this restatement sorts
stably, which is one of the orders ``std::sort`` may produce.
"""
import numpy as np


def _atoll(tok):
    """C atoll: skip leading isspace, optional sign, then decimal digits."""
    i, n = 0, len(tok)
    while i < n and tok[i] in " \t\n\v\f\r":
        i += 1
    sign = 1
    if i < n and tok[i] in "+-":
        sign = -1 if tok[i] == "-" else 1
        i += 1
    v = 0
    while i < n and "0" <= tok[i] <= "9":
        v = v * 10 + (ord(tok[i]) - 48)
        i += 1
    return sign * v


def _split(line):
    """strtok(line, " \\t"): non-empty runs between spaces/tabs."""
    out, cur = [], []
    for ch in line:
        if ch == " " or ch == "\t":
            if cur:
                out.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
    if cur:
        out.append("".join(cur))
    return out


def example_callback(source):
    """The reference example's pre-processing function
    (examples/sssp_basic_single_device.cu:10-26) with a chosen source."""
    def f(argv, s, d, V):
        V[s] = 1073741823 if s != source else 0
        V[d] = 1073741823 if d != source else 0
        if argv:
            w = _atoll(argv[0]) & 0xFFFFFFFF   # atoi, in range for test data
        else:
            w = ((s + 1) * (d + 2) % 127) & 0xFFFFFFFF
        return w, s == source
    return f


def build(text, callback, vpg=1, s_d=True, directed=True, complete=True):
    """Returns a dict with the arrays the reference builds:
    V, R, C, E (in-CSR), frontier (bits per group as produced by the
    callback), frontier_as_reference, dir_R, dir_C, n, m, total_out."""
    V = []
    nbrs = []       # per dst: list of [src, w]
    bits = []       # per group
    for raw in text.split("\n"):          # getline on '\n'; '\r' stays
        if not raw or not ("0" <= raw[0] <= "9"):
            continue
        toks = _split(raw[:255])          # cstrLine[256]
        if len(toks) < 2:
            continue
        a = _atoll(toks[0]) & 0xFFFFFFFF
        b = _atoll(toks[1]) & 0xFFFFFFFF
        argv = toks[2:62]
        mx = max(a, b)
        s, d = (a, b) if s_d else (b, a)
        if len(V) <= mx:
            grow = mx + 1 - len(V)
            V.extend([0] * grow)
            nbrs.extend([] for _ in range(grow))
            ng = -(-(mx + 1) // vpg)
            bits.extend([0] * (ng - len(bits)))
        w, ret = callback(argv, s, d, V)
        nbrs[d].append([s, w])
        bits[d // vpg] |= int(bool(ret))
        if (not directed) and (not complete):
            w2, ret2 = callback(argv, d, s, V)
            nbrs[s].append([d, w2])
            bits[s // vpg] |= int(bool(ret2))

    n0 = len(V)
    out_lists = [[] for _ in range(-(-n0 // vpg))] if directed else []
    for dst, lst in enumerate(nbrs):
        lst.sort(key=lambda x: x[0])
        if directed:
            for s, _ in lst:
                out_lists[s // vpg].append(dst // vpg)
    if directed and vpg != 1:
        out_lists = [sorted(set(x)) for x in out_lists]
    if n0 % 128:
        npad = (n0 // 128 + 1) * 128
        V.extend([0] * (npad - n0))
        nbrs.extend([] for _ in range(npad - n0))
        if directed and vpg == 1:
            out_lists.extend([] for _ in range(npad - n0))
            bits.extend([0] * (npad - n0))
    n = len(V)
    R = np.zeros(n + 1, np.uint32)
    R[1:] = np.cumsum([len(x) for x in nbrs])
    C = np.array([s for lst in nbrs for s, _ in lst], np.uint32)
    E = np.array([w for lst in nbrs for _, w in lst], np.uint32)
    dir_R = np.zeros(len(out_lists) + 1, np.uint32)
    if out_lists:
        dir_R[1:] = np.cumsum([len(x) for x in out_lists])
    dir_C = np.array([x for lst in out_lists for x in lst], np.uint32)
    frontier = np.array(bits, np.uint8)
    as_ref = frontier.copy()
    if not directed:
        as_ref[:] = 0
    return dict(V=np.array(V, np.uint32), R=R, C=C, E=E, frontier=frontier,
                frontier_as_reference=as_ref, dir_R=dir_R, dir_C=dir_C,
                n=n, m=int(R[-1]),
                total_out=int(dir_R[-1]) if (directed and vpg != 1) else int(R[-1]))


def dump_text(g, directed=True):
    """The reference's prepared-graph text dump of a graph built by `build`
    (kites/io/graph_output.cuh:29-35): n, m and 1/0 for directed on lines of
    their own, then V, R, C, E, the frontier bitmap and, for directed graphs,
    dir_R and dir_C, one decimal value per line
    (kites/memory/host_pinned_buffer.cuh:120-124).  The bitmap is the
    graph_csr's vBitmap: ceil(bits / 32) 32-bit words with bit i at position
    i & 31 of word i >> 5 (host_pinned_buffer.cuh:151-153), holding the bits
    graph_csr sets (frontier_as_reference, graph_csr.h:96-104)."""
    bits = g["frontier_as_reference"]
    words = np.zeros((len(bits) + 31) // 32, np.uint64)
    for i in np.flatnonzero(bits):
        words[i >> 5] |= np.uint64(1) << np.uint64(i & 31)
    parts = [f"{g['n']}\n{g['m']}\n{1 if directed else 0}\n"]
    arrays = [g["V"], g["R"], g["C"], g["E"], words]
    if directed:
        arrays += [g["dir_R"], g["dir_C"]]
    for a in arrays:
        parts.append("".join(f"{int(x)}\n" for x in a))
    return "".join(parts)
