"""ORACLE — test infrastructure only (the CPU checker; never imported by the product).

A plain-loop restatement of the integer maps of Muscade's DirectXUA all-steps system:

- asmvec! / asmmat! over the dof groups allΛdofs, allXdofs, allUdofs, allAdofs
  (src/Assemble.jl:321-448, src/DirectXUA.jl:22-56: asm[arrnum(α)], asm[arrnum(α,β)] and the
  class-pair patterns of out.L2);
- makepattern (src/DirectXUA.jl:245-296): which blocks of the big matrix are structurally
  nonzero, a sparse matrix of blocks;
- SparseTools.prepare (src/SparseTools.jl:32-94): the big matrix's colptr/rowval, the block
  assembler (bigmatasm), and the block row/column offsets pgr/pgc (Lvasm/Lvdis);
- sparser! (src/SparseTools.jl:172-196): the drop of small entries before the LU.

Pinned by the goldens of test/TestDirectXUA.jl:109-136 and test/TestSparseTools.jl:36-42,53-62
(tests/test_oracle_direct_maps.py)."""

CLASSES = ("Λ", "X", "U", "A")


def arrnum(a, b=None):
    """arrnum(α) = α, arrnum(α,β) = 4 + β + 4(α−1), 1-based classes (DirectXUA.jl:15-16)."""
    if b is None:
        return a
    return 4 + b + 4 * (a - 1)


def asmvec(index_of_class):
    """asmvec! for the whole-class dof groups (Assemble.jl:340-357): for each element type, the
    element's model dofs of the class (1-based; the group holds every dof of the class, so
    indexedstate is the identity). index_of_class[ieletyp] is a list of per-element dof lists."""
    return [[list(ix) for ix in t] for t in index_of_class]


def asmmat(iasm, jasm, ni, nj):
    """asmmat! (Assemble.jl:373-448): CSC pattern (colptr, rowval, 1-based) of the pairs of
    model dofs, and asm[ieletyp][iele][i + ni_e·(j−1) − 1] = nz number."""
    pairs = []
    for it in range(len(iasm)):
        for e in range(len(iasm[it])):
            ie, je = iasm[it][e], jasm[it][e]
            for jd in range(len(je)):
                for idd in range(len(ie)):
                    if ie[idd] != 0 and je[jd] != 0:
                        pairs.append((je[jd], ie[idd]))
    order = sorted(range(len(pairs)), key=lambda q: pairs[q])   # sortperm of (j,i) tuples
    K = [0] * len(pairs)
    colptr = [0] * (nj + 1)
    colptr[0] = 1
    rowval = []
    inz, icol = 0, 1
    for q in range(len(pairs)):
        if q == 0 or pairs[order[q]] != pairs[order[q - 1]]:
            inz += 1
            j, i = pairs[order[q]]
            rowval.append(i)
            while j > icol:
                icol += 1
                colptr[icol - 1] = inz
        K[order[q]] = inz
    for c in range(icol + 1, nj + 2):
        colptr[c - 1] = inz + 1
    asm, q = [], 0
    for it in range(len(iasm)):
        ta = []
        for e in range(len(iasm[it])):
            ie, je = iasm[it][e], jasm[it][e]
            a = [0] * (len(ie) * len(je))
            for jd in range(len(je)):
                for idd in range(len(ie)):
                    if ie[idd] != 0 and je[jd] != 0:
                        a[idd + len(ie) * jd] = K[q]
                        q += 1
            ta.append(a)
        asm.append(ta)
    return colptr, rowval, asm


def direct_maps(index, ndof, OX, OU, IA):
    """prepare(AssemblyDirect{OX,OU,IA}) (DirectXUA.jl:22-56) on a model described by
    index[c][ieletyp] = per-element model dofs of class c ∈ {X, U, A} and ndof[c]: the asm arrays
    (asm[k] for k = 1..20, per element type, per element) and the class-pair patterns with the
    derivative counts of out.L2 (0×0 for (Λ,Λ))."""
    idx = {"Λ": index["X"], "X": index["X"], "U": index["U"], "A": index["A"]}
    n = {"Λ": ndof["X"], "X": ndof["X"], "U": ndof["U"], "A": ndof["A"]}
    nder = {"Λ": 1, "X": OX + 1, "U": OU + 1, "A": IA}
    asm = {}
    for a, ca in enumerate(CLASSES, 1):
        asm[arrnum(a)] = asmvec(idx[ca])
    pat = {}
    for a, ca in enumerate(CLASSES, 1):
        for b, cb in enumerate(CLASSES, 1):
            cp, rv, am = asmmat(asm[arrnum(a)], asm[arrnum(b)], n[ca], n[cb])
            asm[arrnum(a, b)] = am
            na, nb = (0, 0) if a == b == 1 else (nder[ca], nder[cb])
            pat[(ca, cb)] = (cp, rv, (n[ca], n[cb]), na, nb)
    return asm, pat


def finitediff(order, n, s):
    """(Δs, w) pairs of src/FiniteDifferences.jl:2-34 at step s (1-based) of n."""
    table = [[[(0, 1.)]],
             [[(0, -1.), (1, 1.)], [(-1, -1.), (0, 1.)], [(-1, -.5), (1, .5)]],
             [[(0, 1.), (1, -2.), (2, 1.)], [(-2, 1.), (-1, -2.), (0, 1.)], [(-1, 1.), (0, -2.), (1, 1.)]]]
    if order == 0:
        return table[0][0]
    return table[order][0] if s == 1 else (table[order][1] if s == n else table[order][2])


def makepattern(IA, nstep, pat):
    """makepattern (DirectXUA.jl:245-296): the nonzero blocks (block row, block column, class
    pair), unique in first-appearance order, as a CSC matrix of blocks: (colptr, rowval, blocks)
    with blocks[k] = the class pair whose pattern block k carries."""
    entries = []
    cum = 0
    for ns in nstep:
        for istep in range(1, ns + 1):
            for a, ca in enumerate(CLASSES[:3], 1):
                for b, cb in enumerate(CLASSES[:3], 1):
                    _, _, _, na, nb = pat[(ca, cb)]
                    for ad in range(1, na + 1):
                        for bd in range(1, nb + 1):
                            for (da, _) in finitediff(ad - 1, ns, istep):
                                for (db, _) in finitediff(bd - 1, ns, istep):
                                    entries.append((cum + 3 * (istep + da - 1) + a, cum + 3 * (istep + db - 1) + b,
                                                    (ca, cb)))
        cum += 3 * ns
    if IA == 1:
        Ablk = 3 * sum(nstep) + 1
        entries.append((Ablk, Ablk, ("A", "A")))
        cum = 0
        for ns in nstep:
            for istep in range(1, ns + 1):
                for a, ca in enumerate(CLASSES[:3], 1):
                    if pat[("A", ca)][3] > 0:
                        entries.append((Ablk, cum + 3 * (istep - 1) + a, ("A", ca)))
                        entries.append((cum + 3 * (istep - 1) + a, Ablk, (ca, "A")))
            cum += 3 * ns
    seen, uniq = set(), []
    for r, c, p in entries:
        if (r, c) not in seen:
            seen.add((r, c))
            uniq.append((r, c, p))
    nb = 3 * sum(nstep) + (1 if IA == 1 else 0)
    cols = [[] for _ in range(nb)]
    for r, c, p in uniq:
        cols[c - 1].append((r, p))
    colptr, rowval, blocks = [1], [], []
    for c in range(nb):
        for r, p in sorted(cols[c]):
            rowval.append(r)
            blocks.append(p)
        colptr.append(len(rowval) + 1)
    return colptr, rowval, blocks


def sparse_prepare(bcolptr, browval, blocks, pat):
    """SparseTools.prepare (SparseTools.jl:32-94) for a pattern of class-pair blocks: the big
    matrix's colptr, rowval (1-based), the assembler asm_igv[ibv] (1-based global value number of
    each value of block ibv) and pgr, pgc (Lvasm, Lvdis)."""
    nbc = len(bcolptr) - 1
    nbr = max(browval) if browval else 0
    nlr = [-1] * (nbr + 1)
    nlc = [-1] * (nbc + 1)
    for ibc in range(1, nbc + 1):
        for ibv in range(bcolptr[ibc - 1], bcolptr[ibc]):
            ibr = browval[ibv - 1]
            m, n = pat[blocks[ibv - 1]][2]
            if nlr[ibr] not in (-1, m) or nlc[ibc] not in (-1, n):
                raise ValueError("blocks of one block row (column) differ in size")
            nlr[ibr], nlc[ibc] = m, n
    nlr[0] = nlc[0] = 1
    if -1 in nlr or -1 in nlc:
        raise ValueError("invalid sparse-of-sparse pattern")
    pgr, pgc = [], []
    acc = 0
    for v in nlr:
        acc += v
        pgr.append(acc)
    acc = 0
    for v in nlc:
        acc += v
        pgc.append(acc)
    ngc = pgc[-1] - 1
    pigr = [1] + [0] * ngc
    igr = []
    asm_igv = [[0] * len(pat[blocks[k]][1]) for k in range(len(blocks))]
    igv = 1
    for ibc in range(1, nbc + 1):
        for ilc in range(1, nlc[ibc] + 1):
            igc = pgc[ibc - 1] - 1 + ilc
            for ibv in range(bcolptr[ibc - 1], bcolptr[ibc]):
                ibr = browval[ibv - 1]
                cp, rv, _, _, _ = pat[blocks[ibv - 1]]
                for ilv in range(cp[ilc - 1], cp[ilc]):
                    igr.append(pgr[ibr - 1] - 1 + rv[ilv - 1])
                    asm_igv[ibv - 1][ilv - 1] = igv
                    igv += 1
            pigr[igc] = igv
    return pigr, igr, asm_igv, pgr, pgc


def sparser(colptr, rowval, nzval, rtol=1e-9):
    """sparser!(S, rtol) (SparseTools.jl:172-196): keep |v| ≥ rtol·max|v|."""
    atol = rtol * max((abs(v) for v in nzval), default=0.)
    cp, rv, nz = [1], [], []
    for c in range(len(colptr) - 1):
        for k in range(colptr[c] - 1, colptr[c + 1] - 1):
            if abs(nzval[k]) >= atol:
                rv.append(rowval[k])
                nz.append(nzval[k])
        cp.append(len(rv) + 1)
    return cp, rv, nz
