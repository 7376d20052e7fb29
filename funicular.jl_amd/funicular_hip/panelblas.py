"""Panel BLAS: gram, panelmul, cholqr2, rightmul, project and the
elementwise scale / axpy / dot / norm (src/panelblas.jl:1-370), on the sweep
pipeline and the HIP kernels.

Under the default "reference" schedule the steps and copies are the
reference's: gram and the triangular solve sweep row blocks of height
floor(N w / k); cholqr2 is two passes, and a third when a shifted Cholesky
fired, each pass a Gram sweep, the k x k Gram down to the host, its Cholesky
in LAPACK (src/panelblas.jl:148-155), the factor up and a solve sweep; project
runs p (p + 1) panel steps labelled (applying, i, j). The "fused" schedule
(a plan's `schedule=` or the `schedule=` keyword) shortens them as the
functions below describe; its results agree to rounding.
"""
import numpy as np
import scipy.linalg

from . import plan as _plan
from .device_api import ArgumentError
from .executor import PanelSteps, RowSteps, readpanels, rowtraversal, sweep, updatepanels, writepanels
from .operator import apply_operator, assert_operator_shape, assert_square_operator
from .panels import (assert_conformal, assert_same_panels, assert_same_rows, host_storage_eltype, npanels, panelrange,
                     panelwidth, pooled_like)


def resolve_schedule(pm, schedule=None):
    """The schedule an operation on pm runs: the keyword when given, else the
    plan's ("reference" unless the plan says otherwise)."""
    s = pm.plan.schedule if schedule is None else schedule
    if s not in _plan.SCHEDULES:
        raise ArgumentError(f"schedule must be one of {', '.join(map(repr, _plan.SCHEDULES))}, got {s!r}")
    return s


def shared_plan(X, Y):
    """src/panelblas.jl:4-8."""
    if X.dtype != Y.dtype:
        raise ArgumentError(f"the two panel matrices compute in {X.dtype} and {Y.dtype}; an operation takes one "
                            "compute eltype, and a narrower eltype is set on the plan's storage tiers instead")
    if X.plan is not Y.plan:
        raise ArgumentError("the two panel matrices come from different ResidencyPlans; an operation stages "
                            "through one plan's buffers, so build both from a single plan")
    return X.plan


def carry(step):
    """The first step overwrites the accumulator, later ones add (src/panelblas.jl:55-59)."""
    return 0.0 if step == 0 else 1.0


def collect_small(device, plan):
    """src/panelblas.jl:61-65."""
    host = np.empty(device.shape, dtype=device.dtype, order="F")
    plan.backend.d2h(host, device)
    return host


def gram(X, Y=None, *, nbuffers=None, compute=None):
    """The k x k matrix X^H Y: each row block adds its product to a device
    accumulator, which comes to the host at the end (src/panelblas.jl:10-53). `compute=np.complex64`
    forms it on the FP32 matrix cores with FP64 sums across row chunks."""
    Y = X if Y is None else Y
    plan = shared_plan(X, Y)
    assert_conformal(X, Y)
    accumulator = _plan.checkout_device_buffers(plan, X.dtype, (X.k, X.k), 1)[0]
    try:
        gram_into(accumulator, X, Y, nbuffers=nbuffers, compute=compute)
        return collect_small(accumulator, plan)
    finally:
        _plan.checkin_device_buffers(plan, [accumulator])


def gram_into(accumulator, X, Y, *, nbuffers=None, compute=None):
    """gram's row sweep into a k x k device accumulator, left on the device."""
    same = Y is X
    plan = X.plan
    T = X.dtype
    backend = plan.backend
    traversal = rowtraversal(X)
    height = traversal.blocks[0][1] - traversal.blocks[0][0]
    work = _plan.checkout_workspace(
        plan, backend.gram_workspace_bytes(height, X.k, X.k, T, hermitian=same, compute=compute))
    try:
        if same:
            def step(block, Xb):
                backend.gram_accumulate(accumulator, Xb, Xb, 1.0, carry(block), work, compute)
            sweep(step, readpanels(X), nbuffers=nbuffers, traversal=traversal)
        else:
            def step(block, Xb, Yb):
                backend.gram_accumulate(accumulator, Xb, Yb, 1.0, carry(block), work, compute)
            sweep(step, readpanels(X), readpanels(Y), nbuffers=nbuffers, traversal=traversal)
    finally:
        _plan.checkin_device_buffers(plan, [work])
    return accumulator


def panelmul(Y, G, X, *, nbuffers=None):
    """Y_j <- G X_j for every column panel j (src/panelblas.jl:67-100)."""
    shared_plan(X, Y)
    if X is Y:
        raise ArgumentError("panelmul needs distinct input and output matrices: applied panel by panel in "
                            "place, the operator would meet columns that hold its own earlier output")
    assert_same_panels(X, Y)
    assert_operator_shape(G, Y.N, X.N)
    sweep(lambda _, Xb, Yb: apply_operator(Yb, G, Xb), readpanels(X), writepanels(Y), nbuffers=nbuffers)
    return Y


def cholqr2(Y, *, nbuffers=None, schedule=None):
    """Orthonormalizes Y in place by CholeskyQR2 and returns the k x k upper
    triangular R for which Q R reproduces the input (src/panelblas.jl:102-139).
    A pass whose Gram needed a shift triggers a third pass (shifted
    CholeskyQR3, Fukaya et al.).

    schedule="reference" runs the reference's four row sweeps (six with the
    third pass) with the Cholesky on the host; "fused" keeps G and R on the
    device and, off HBM, forms each next pass's Gram inside the solve sweep,
    so Y goes up three times instead of four."""
    if resolve_schedule(Y, schedule) == "reference":
        return _reference_passes(Y, nbuffers)
    return _passes(Y, nbuffers)


def triangular_product(A, B):
    """A B for upper triangular A and B (BLAS trmm: half the flops of a full
    product; the result is upper triangular)."""
    A = np.asfortranarray(A)
    B = np.asfortranarray(B, dtype=A.dtype)
    trmm, = scipy.linalg.get_blas_funcs(("trmm",), (A, B))
    return trmm(1.0, A, B, side=0, lower=0, trans_a=0, diag=0)


_product_pool = None


def _device_pass(Y, nbuffers, R_prev=None, first=None, gram_ready=None, fuse_next=None):
    """One CholQR pass with G, R and the solve's factor kept on the device
    (src/panelblas.jl:141-146): the Gram sweep into a device accumulator, its
    plain Cholesky factorization on the device, the solve sweep from the
    device factor. Only R comes back to the host (4 MiB at k = 512), where the
    product R R_prev is formed on a thread while the solve runs. When the
    device factorization meets a bad pivot, G goes to the host and the
    reference's rule takes over there (a plain retry, then the shift).

    first=(shift, solve) makes it the mixed first pass of cholqr2_mixed: FP32
    Gram and Cholesky, the solve in FP32 or FP64.

    gram_ready is a device accumulator that already holds this pass's Gram
    (the previous pass's solve sweep formed it); the pass takes it over.
    fuse_next(shifted) says, once this pass's Cholesky is known, whether its
    solve sweep should form the next pass's Gram as well, which saves that
    pass a full sweep over Y (a full upload of it when Y lives in host
    memory). Returns (R_total, shifted, next_gram) for an FP64 pass and
    (R, s, next_gram) for the mixed one."""
    global _product_pool
    plan = Y.plan
    backend = plan.backend
    T = Y.dtype
    k = Y.k
    acc = gram_ready
    factor = info = narrow = nxt = work_g = rounding = None
    try:
        if acc is None:
            acc = _plan.checkout_device_buffers(plan, T, (k, k), 1)[0]
        factor = _plan.checkout_device_buffers(plan, T, (k, k), 1)[0]
        info = _plan.checkout_device_buffers(plan, np.int32, (1, 1), 1)[0]
        compute = None if first is None else np.complex64
        if gram_ready is None:
            gram_into(acc, Y, Y, nbuffers=nbuffers, compute=compute)
        shifted, s = False, 0.0
        bad = 1
        if first is None:
            backend.potrf_upper(factor, acc, info)
            bad = int(backend.to_host(info)[0, 0])
        elif first[0] == "adaptive":
            narrow = _plan.checkout_device_buffers(plan, np.complex64, (k, k), 1)[0]
            backend.potrf_upper(narrow, acc, info)
            bad = int(backend.to_host(info)[0, 0])
            if bad == 0:
                backend.convert_copy(factor, narrow)
        if bad == 0:
            R = backend.to_host(factor)
        else:
            G = collect_small(acc, plan)
            if first is None:
                R, shifted = cholesky_upper(G, Y.N, k)
            else:
                R32, s = cholesky_upper_fp32(G, Y.N, k, first[0])
                R = R32.astype(T)
            backend.h2d(factor, np.asfortranarray(R.astype(T, copy=False)))
        product = None
        if R_prev is not None:
            if _product_pool is None:
                import concurrent.futures
                _product_pool = concurrent.futures.ThreadPoolExecutor(max_workers=1,
                                                                      thread_name_prefix="cholqr-product")
            product = _product_pool.submit(triangular_product, R, R_prev)
        solve = None if first is None or first[1] == "fp64" else np.complex64
        rounding = None
        if fuse_next is not None and fuse_next(shifted or s > 0.0):
            traversal = rowtraversal(Y)
            height = traversal.blocks[0][1] - traversal.blocks[0][0]
            nxt = _plan.checkout_device_buffers(plan, T, (k, k), 1)[0]
            work_g = _plan.checkout_workspace(plan, backend.gram_workspace_bytes(height, k, k, T, hermitian=True))
            narrow_to = _fusable(Y)[1]
            if narrow_to is not None:
                rounding = _plan.checkout_device_buffers(plan, narrow_to, (height, k), 1)[0]
        try:
            rdiv_rows(Y, R, nbuffers=nbuffers, compute=solve, factor=factor,
                      next_gram=None if nxt is None else (nxt, work_g, rounding), schedule="fused")
        finally:
            total = R if product is None else product.result()
        handed, nxt = nxt, None
        return ((total, shifted) if first is None else (R, s)) + (handed,)
    finally:
        _plan.checkin_device_buffers(plan, [acc, factor, info, narrow, nxt, work_g, rounding])


def cholqr_pass_then(Y, R_prev, nbuffers=None):
    """One more CholQR pass over Y on the fused schedule, returning
    (R_pass R_prev, shifted). The host forms the triangular product while the
    pass's solve sweep runs on the device, instead of after it."""
    return _device_pass(Y, nbuffers, R_prev=R_prev)[:2]


def cholqr_pass(Y, nbuffers=None, schedule=None):
    """One CholQR pass, returning (R, shifted) (src/panelblas.jl:141-146):
    the Gram sweep, the Cholesky (on the host under the reference schedule,
    on the device under the fused one) and the solve sweep."""
    if resolve_schedule(Y, schedule) == "reference":
        G = gram(Y, nbuffers=nbuffers)
        R, shifted = cholesky_upper(G, Y.N, Y.k)
        rdiv_rows(Y, R, nbuffers=nbuffers, schedule="reference")
        return R, shifted
    return _device_pass(Y, nbuffers)[:2]


def _reference_passes(Y, nbuffers, first=None):
    """cholqr2 (first=None) or cholqr2_mixed on the reference's schedule
    (src/panelblas.jl:130-146): every pass is its own Gram sweep, host
    Cholesky and solve sweep. first=(shift, solve) makes the first pass the
    mixed one: its Gram on the FP32 matrix cores with FP64 sums, its
    Cholesky in FP32 on the host, the solve in FP32 or FP64."""
    if first is None:
        R, shifted1 = cholqr_pass(Y, nbuffers, schedule="reference")
    else:
        shift, solve = first
        G = gram(Y, nbuffers=nbuffers, compute=np.complex64)
        R32, s = cholesky_upper_fp32(G, Y.N, Y.k, shift)
        R = R32.astype(Y.dtype)
        rdiv_rows(Y, R, nbuffers=nbuffers, compute=None if solve == "fp64" else np.complex64, schedule="reference")
        shifted1 = s > 0.0
    R2, shifted2 = cholqr_pass(Y, nbuffers, schedule="reference")
    R = triangular_product(R2, R)
    if shifted1 or shifted2:
        R3, _ = cholqr_pass(Y, nbuffers, schedule="reference")
        R = triangular_product(R3, R)
    return R


def _fusable(Y):
    """Whether the solve sweeps of Y may form the next pass's Gram, and the
    eltype a solved block is rounded to first: that of the host tier when it
    narrows, so the Gram sees exactly the values the next pass would read
    back; None when nothing narrows. A file or disk tier that narrows on its
    own rules fusion out."""
    plan = Y.plan
    T = np.dtype(Y.dtype)
    store = getattr(Y, "store", None)
    if plan.disk_eltype is not None and np.dtype(plan.disk_eltype) != T:
        return False, None
    if store is not None and np.dtype(store.stored) != T:
        return False, None
    if plan.host_eltype is not None and np.dtype(plan.host_eltype) != T:
        return True, np.dtype(plan.host_eltype)
    return True, None


def _passes(Y, nbuffers, first=None):
    """The passes of cholqr2 (first=None) or cholqr2_mixed: the first pass,
    the FP64 pass, and a third FP64 pass when a Cholesky needed a shift
    (src/panelblas.jl:130-139). Off HBM each solve sweep also forms the next
    pass's Gram (the same products in the same order, so the same bits), and
    only when a next pass will run: after pass two, that is known from the two
    Cholesky factorizations before its solve starts."""
    fusable, _ = _fusable(Y)
    fuse = not Y.resident and fusable
    R, shift1, g = _device_pass(Y, nbuffers, first=first, fuse_next=(lambda _: True) if fuse else None)
    shifted1 = shift1 > 0.0 if first is not None else shift1
    # A pass given gram_ready owns it from then on, and checks it back in.
    R, shifted2, g2 = _device_pass(Y, nbuffers, R_prev=R, gram_ready=g,
                                   fuse_next=(lambda s2: shifted1 or s2) if fuse else None)
    if shifted1 or shifted2:
        R, _, _ = _device_pass(Y, nbuffers, R_prev=R, gram_ready=g2)
    return R


def _potrf(G):
    """Upper Cholesky factor of the Hermitian matrix whose upper triangle G
    holds (LAPACK potrf, as cholesky(Hermitian(G)) does); None if it fails."""
    potrf, = scipy.linalg.get_lapack_funcs(("potrf",), (G,))
    # clean=True zeroes the strictly lower triangle of the factor.
    c, info = potrf(G, lower=False, clean=True, overwrite_a=False)
    if info != 0:
        return None
    return np.asfortranarray(c)


def cholesky_upper(G, N, k):
    """The upper Cholesky factor of G, retried with G + s I,
    s = 11 N k eps ||G||_F, when G is not numerically positive definite
    (src/panelblas.jl:148-155). Returns (R, shifted)."""
    G = np.asarray(G)
    plain = _potrf(G)
    if plain is not None:
        return plain, False
    eps = np.finfo(G.dtype).eps
    shift = 11 * N * k * eps * np.linalg.norm(G)
    shifted = _potrf(G + shift * np.eye(G.shape[0], dtype=G.dtype))
    if shifted is None:
        raise ArgumentError(
            f"the Cholesky factorization of the {k}×{k} Gram failed again with the diagonal raised by {shift}: "
            f"the columns are dependent to working precision (a zero or repeated column, or a condition number "
            f"beyond {1 / eps}). Remove the dependent columns and orthonormalize the rest")
    return shifted, True


def rdiv_rows(Y, R, *, nbuffers=None, compute=None, factor=None, next_gram=None, schedule=None):
    """Y <- Y R^-1 over row blocks (src/panelblas.jl:157-169), R uploaded
    first. Under the fused schedule a matrix resident in HBM is solved in one
    launch over all its rows instead of block by block. The keywords below
    serve the fused cholqr2. `factor`, a device copy of R already in place,
    saves the upload. next_gram=(acc,
    work) also accumulates the Gram of the solved blocks into the device
    matrix acc while each block is still on the device: the next CholQR
    pass's Gram, without a sweep of its own; next_gram=(acc, work, rounding)
    with a device buffer `rounding` of the host tier's narrower eltype first
    rounds each solved block to it, as storing it would."""
    plan = Y.plan
    T = Y.dtype
    backend = plan.backend
    owned = factor is None
    if owned:
        factor = _plan.checkout_device_buffers(plan, T, R.shape, 1)[0]
    work = None
    try:
        work = _plan.checkout_workspace(plan, backend.trsm_workspace_bytes(Y.k, T if compute is None else compute))
        if owned:
            backend.h2d(factor, np.asfortranarray(R.astype(T, copy=False)))
        # The first block folds the factor into the workspace; the later ones,
        # queued behind it on the compute stream, reuse it.
        def step(block, Yb):
            if block == 0:
                backend.rdiv_upper(Yb, factor, work, compute)
            else:
                backend.rdiv_upper_folded(Yb, work, compute)
            if next_gram is not None:
                acc, gwork, rounding = next_gram
                if rounding is not None:
                    # Round the solved block to the host eltype it is stored in,
                    # so the Gram is that of the values the next pass reads.
                    r = rounding.view((0, Yb.rows), None)
                    backend.convert_copy(r, Yb)
                    backend.convert_copy(Yb, r)
                backend.gram_accumulate(acc, Yb, Yb, 1.0, carry(block), gwork)
        # Rows are independent, so a matrix resident in HBM needs no row
        # blocks: one launch over all N rows leaves no partly filled last wave
        # of workgroups per block (3% of a 629,856-row launch). A Gram rides
        # along only on row blocks, its chunking being that of gram's sweep.
        whole = Y.resident and next_gram is None and resolve_schedule(Y, schedule) == "fused"
        traversal = RowSteps(Y.N, Y.N) if whole else rowtraversal(Y)
        sweep(step, updatepanels(Y), nbuffers=nbuffers, traversal=traversal)
    finally:
        _plan.checkin_device_buffers(plan, [factor, work] if owned else [work])
    return Y


def rightmul(*args, nbuffers=None):
    """rightmul(Y, C): Y <- Y C in place (C k x k); rightmul(dest, src, C):
    dest <- src C (src/panelblas.jl:171-245)."""
    if len(args) == 2:
        return _rightmul_inplace(args[0], args[1], nbuffers)
    if len(args) == 3:
        return _rightmul_into(args[0], args[1], args[2], nbuffers)
    raise TypeError("rightmul takes (Y, C) or (dest, src, C)")


def _rightmul_inplace(Y, C, nbuffers):
    C = np.asarray(C)
    if C.shape != (Y.k, Y.k):
        raise ArgumentError(f"rightmul(Y, C) overwrites Y, whose {Y.k} columns are fixed, so C has to be "
                            f"{Y.k}×{Y.k}; it is {C.shape}. To change the number of columns, write the product into "
                            "another matrix with rightmul(dest, Y, C)")
    from .panels import assert_writable
    assert_writable(Y)
    plan = Y.plan
    T = Y.dtype
    backend = plan.backend
    traversal = rowtraversal(Y)
    from .executor import buffer_dims
    factor = _plan.checkout_device_buffers(plan, T, (Y.k, Y.k), 1)[0]
    scratch = _plan.checkout_device_buffers(plan, T, buffer_dims(traversal, Y), 1)[0]
    try:
        backend.h2d(factor, np.asfortranarray(C.astype(T)))

        def step(_, Yb):
            # The last block may be shorter than the buffer (src/panelblas.jl:212-216).
            product = scratch.view((0, Yb.rows), None)
            backend.rightmul_gemm(product, Yb, factor)
            backend.convert_copy(Yb, product)
        sweep(step, updatepanels(Y), nbuffers=nbuffers, traversal=traversal)
    finally:
        _plan.checkin_device_buffers(plan, [factor, scratch])
    return Y


def _rightmul_into(dest, src, C, nbuffers):
    if dest is src:
        raise ArgumentError("rightmul(dest, src, C) reads src while it writes dest, so they must be two matrices; "
                            "for a product that replaces its own input call rightmul(Y, C)")
    C = np.asarray(C)
    if C.shape != (src.k, dest.k):
        raise ArgumentError(f"rightmul(dest, src, C) with a {src.N}×{src.k} src and a {dest.N}×{dest.k} dest "
                            f"needs C to be {src.k}×{dest.k}; it is {C.shape}")
    plan = shared_plan(src, dest)
    assert_same_rows(src, dest)
    T = dest.dtype
    backend = plan.backend
    # The shorter of the two row-block heights serves both (src/panelblas.jl:230-234).
    height = min(_plan.row_block_height(src.N, src.k, src.w), _plan.row_block_height(dest.N, dest.k, dest.w))
    traversal = RowSteps(src.N, height)
    factor = _plan.checkout_device_buffers(plan, T, (src.k, dest.k), 1)[0]
    try:
        backend.h2d(factor, np.asfortranarray(C.astype(T)))
        sweep(lambda _, sb, db: backend.rightmul_gemm(db, sb, factor), readpanels(src), writepanels(dest),
              nbuffers=nbuffers, traversal=traversal)
    finally:
        _plan.checkin_device_buffers(plan, [factor])
    return dest


def project_schedule(p):
    """The panels and labels of project's sweep: for each j, apply G to panel j,
    then reduce every panel i against it (src/panelblas.jl:272-281)."""
    panels, labels = [], []
    for j in range(p):
        panels.append(j)
        labels.append((True, j, j))
        for i in range(p):
            panels.append(i)
            labels.append((False, i, j))
    return panels, labels


def _pool_can_hold(plan, T, dims):
    pool = plan.devicepool
    if pool.available.get((np.dtype(T), (int(dims[0]), int(dims[1])))):
        return True
    return pool.allocated + int(dims[0]) * int(dims[1]) * np.dtype(T).itemsize <= pool.budget


def project(Q, G, *, nbuffers=None, group=None, schedule=None):
    """S = Q^H G Q without storing G Q (src/panelblas.jl:247-301).

    Under the reference schedule (the default) this is the reference's
    p (p + 1) panel steps, labelled (applying, i, j), whatever tier Q is on;
    `group` > 1 asks for runs of held images on any schedule. The fused
    schedule picks its own route:

    A basis resident in HBM whose plan can lend a second N x k buffer takes a
    shorter route, since no panel has to travel: G Q is formed whole in that
    pooled buffer, one stencil pass per panel, and a single Gram sweep gives
    all of S on 64 x 64 matrix-core tiles, instead of p^2 reductions of 32 x 32
    tiles, each reading two whole panels from HBM.

    Otherwise the images of a run of g consecutive panels are held on the
    device (g = `group`, by default as many as the plan's device pool can lend
    beside one sweep's staging buffers): the run's panels are staged once to
    form G Q_run, then every panel Q_i streams past it once for the blocks
    Q_i^H (G Q_run) of S. That is p + p ceil(p / g) panel transfers instead of
    the reference schedule's p (p + 1); with g = 1 the reference schedule runs
    as it is, its labels included."""
    plan = Q.plan
    assert_square_operator(G, Q.N)
    if group is None and resolve_schedule(Q, schedule) == "reference":
        group = 1
    if group is None and Q.resident and _pool_can_hold(plan, Q.dtype, (Q.N, Q.k)):
        with pooled_like(Q) as GQ:
            panelmul(GQ, G, Q, nbuffers=nbuffers)
            return gram(Q, GQ, nbuffers=nbuffers)
    g = project_group(Q, nbuffers) if group is None else max(1, int(group))
    if g > 1:
        return _project_runs(Q, G, g, nbuffers)
    p = npanels(Q)
    T = Q.dtype
    backend = plan.backend
    panels, labels = project_schedule(p)
    image = _plan.checkout_device_buffers(plan, T, (Q.N, Q.w), 1)[0]
    accumulator = None
    work = None
    try:
        accumulator = _plan.checkout_device_buffers(plan, T, (Q.k, Q.k), 1)[0]
        work = _plan.checkout_workspace(plan, backend.gram_workspace_bytes(Q.N, Q.w, Q.w, T, hermitian=False))

        def step(label, Qb):
            applying, i, j = label
            GQj = image.view(None, (0, panelwidth(Q, j)))
            if applying:
                apply_operator(GQj, G, Qb)
            else:
                ri, rj = panelrange(Q, i), panelrange(Q, j)
                block = accumulator.view((ri.start, ri.stop), (rj.start, rj.stop))
                backend.gram_accumulate(block, Qb, GQj, 1.0, 0.0, work)
        sweep(step, readpanels(Q), nbuffers=nbuffers, traversal=PanelSteps(panels, labels))
        return collect_small(accumulator, plan)
    finally:
        _plan.checkin_device_buffers(plan, [image, accumulator, work])


def _project_runs_bytes(Q, g):
    """Device bytes of project's run route with runs of g panels: the run's
    images, the k x k accumulator and the largest Gram workspace."""
    item = np.dtype(Q.dtype).itemsize
    spans = sorted({c1 - c0 for _, _, c0, c1 in _groups(Q, g)})
    widths = sorted({panelwidth(Q, i) for i in range(npanels(Q))})
    backend = Q.plan.backend
    work = max(backend.gram_workspace_bytes(Q.N, wi, ws, Q.dtype, hermitian=False) for wi in widths for ws in spans)
    return Q.N * spans[-1] * item + Q.k * Q.k * item + work


def project_group(Q, nbuffers=None):
    """How many panel images project holds on the device at once: the longest
    run whose images, accumulator and Gram workspace the plan's device pool can
    still lend beside one sweep's staging buffers (those it does not hold
    already), at most every panel, at least one."""
    plan = Q.plan
    pool = plan.devicepool
    panel = Q.N * Q.w * np.dtype(Q.dtype).itemsize
    spare = pool.budget - pool.allocated - _panel_staging_missing(Q, nbuffers)
    g = min(npanels(Q), max(1, spare // max(panel, 1)))
    while g > 1 and _project_runs_bytes(Q, g) > spare:
        g -= 1
    return int(g)


def _project_runs(Q, G, g, nbuffers):
    """project with the images of g consecutive panels held on the device."""
    plan = Q.plan
    backend = plan.backend
    T = Q.dtype
    p = npanels(Q)
    runs = list(_groups(Q, g))
    widths = sorted({panelwidth(Q, i) for i in range(p)})
    spans = sorted({c1 - c0 for _, _, c0, c1 in runs})
    image = _plan.checkout_device_buffers(plan, T, (Q.N, spans[-1]), 1)[0]
    accumulator = work = None
    try:
        accumulator = _plan.checkout_device_buffers(plan, T, (Q.k, Q.k), 1)[0]
        need = max(backend.gram_workspace_bytes(Q.N, wi, ws, T, hermitian=False) for wi in widths for ws in spans)
        work = _plan.checkout_workspace(plan, need)
        for j0, j1, c0, c1 in runs:
            def apply(j, Qb, c0=c0):
                r = panelrange(Q, j)
                apply_operator(image.view(None, (r.start - c0, r.stop - c0)), G, Qb)
            sweep(apply, readpanels(Q), nbuffers=nbuffers, traversal=PanelSteps(list(range(j0, j1))))
            GQrun = image.view(None, (0, c1 - c0))

            def reduce(i, Qb, c0=c0, c1=c1, GQrun=GQrun):
                ri = panelrange(Q, i)
                backend.gram_accumulate(accumulator.view((ri.start, ri.stop), (c0, c1)), Qb, GQrun, 1.0, 0.0, work)
            sweep(reduce, readpanels(Q), nbuffers=nbuffers, traversal=PanelSteps(list(range(p))))
        return collect_small(accumulator, plan)
    finally:
        _plan.checkin_device_buffers(plan, [image, accumulator, work])


# CholeskyQR by column panels, for a basis that the host budget cannot hold
# whole. A row sweep stages every panel at once (src/executor.jl:66-67), so
# gram and cholqr2 refuse such a matrix; these sweeps hold a run of panels on
# the device while the others stream past it, so only the run and the
# staging buffers need to fit. The reference has no such traversal.

def _panel_staging_missing(X, nbuffers):
    """Bytes of the staging buffers a column-panel sweep over X needs (nb
    panels, and nb more in the host storage eltype for a narrowed host tier)
    that the plan's device pool does not already hold idle."""
    if X.resident:
        return 0
    plan = X.plan
    nb = plan.nbuffers if nbuffers is None else int(nbuffers)
    dims = (int(X.N), int(X.w))
    kinds = [np.dtype(X.dtype)]
    S = np.dtype(host_storage_eltype(plan, X.dtype))
    if S != kinds[0]:
        kinds.append(S)
    missing = 0
    for T in kinds:
        idle = len(plan.devicepool.available.get((T, dims), ()))
        missing += max(0, nb - idle) * dims[0] * dims[1] * T.itemsize
    return missing


def panel_group(X, nbuffers=None):
    """How many consecutive column panels the panel-wise sweeps hold on the
    device at once: as many as the plan's device pool can still lend beside
    one sweep's staging buffers (those it does not hold already) and the
    kernels' workspaces, and at least one."""
    plan = X.plan
    pool = plan.devicepool
    panel = X.N * X.w * np.dtype(X.dtype).itemsize
    spare = pool.budget - pool.allocated - _panel_staging_missing(X, nbuffers) - 2 * panel - (64 << 20)
    return int(max(1, min(npanels(X), spare // max(panel, 1))))


def _groups(X, g):
    """(j0, j1, c0, c1) for each run of g consecutive panels: panels j0..j1-1,
    columns c0..c1-1."""
    p = npanels(X)
    for j0 in range(0, p, g):
        j1 = min(p, j0 + g)
        yield j0, j1, panelrange(X, j0).start, panelrange(X, j1 - 1).stop


def _hold_group(X, T, gcols):
    plan = X.plan
    return _plan.checkout_device_buffers(plan, T, (X.N, gcols), 1)[0]


def gram_panelwise(X, *, nbuffers=None, compute=None, group=None):
    """The Hermitian k x k Gram X^H X by column panels, for a matrix the row
    sweeps cannot stage whole. A run of g consecutive panels (g = `group`, by
    default panel_group's) is copied into one device buffer and its own block
    of G is one Hermitian Gram; then every earlier panel streams past it once
    for its blocks against the run. Each panel is read 1 + (number of runs
    after it) times: p (p + 1) / 2 reads in all for g = 1, p for g = p. The
    lower triangle is mirrored on the host. compute=np.complex64 forms the
    blocks on the FP32 matrix cores, as gram does."""
    plan = X.plan
    backend = plan.backend
    T = X.dtype
    g = panel_group(X, nbuffers) if group is None else max(1, int(group))
    gmax = max(c1 - c0 for _, _, c0, c1 in _groups(X, g))
    hold = _hold_group(X, T, gmax)
    acc = work = None
    try:
        acc = _plan.checkout_device_buffers(plan, T, (X.k, X.k), 1)[0]
        need = max(backend.gram_workspace_bytes(X.N, gmax, gmax, T, hermitian=True, compute=compute),
                   backend.gram_workspace_bytes(X.N, X.w, gmax, T, hermitian=False, compute=compute))
        work = _plan.checkout_workspace(plan, need)
        for j0, j1, c0, c1 in _groups(X, g):
            held = hold.view(None, (0, c1 - c0))

            # The run's panels into the device buffer; its own block once the
            # last one is in (same compute stream).
            def load(j, Xb, j1=j1, c0=c0, c1=c1, held=held):
                r = panelrange(X, j)
                backend.convert_copy(held.view(None, (r.start - c0, r.stop - c0)), Xb)
                if j == j1 - 1:
                    backend.gram_accumulate(acc.view((c0, c1), (c0, c1)), held, held, 1.0, 0.0, work, compute)
            sweep(load, readpanels(X), nbuffers=nbuffers, traversal=PanelSteps(list(range(j0, j1))))
            if j0 > 0:
                def cross(i, Xb, c0=c0, c1=c1, held=held):
                    r = panelrange(X, i)
                    backend.gram_accumulate(acc.view((r.start, r.stop), (c0, c1)), Xb, held, 1.0, 0.0, work, compute)
                sweep(cross, readpanels(X), nbuffers=nbuffers, traversal=PanelSteps(list(range(j0))))
        G = collect_small(acc, plan)
    finally:
        _plan.checkin_device_buffers(plan, [hold, acc, work])
    return np.triu(G) + np.triu(G, 1).conj().T


def rdiv_panelwise(Y, R, *, nbuffers=None, group=None):
    """Y <- Y R^-1 by column panels, left looking over runs of g consecutive
    panels (g = `group`, by default panel_group's): a run is copied into one
    device buffer, every solved panel before it streams past once
    (Y_run <- Y_run - Q_i R_i,run, on the scaled row-block product), the run
    is solved by its own diagonal block of R (the solve kernel over the run's
    columns), and its panels are written back."""
    plan = Y.plan
    backend = plan.backend
    T = Y.dtype
    g = panel_group(Y, nbuffers) if group is None else max(1, int(group))
    gmax = max(c1 - c0 for _, _, c0, c1 in _groups(Y, g))
    hold = _hold_group(Y, T, gmax)
    factor = work = None
    try:
        factor = _plan.checkout_device_buffers(plan, T, R.shape, 1)[0]
        backend.h2d(factor, np.asfortranarray(R.astype(T, copy=False)))
        work = _plan.checkout_workspace(plan, backend.trsm_workspace_bytes(gmax, T))
        for j0, j1, c0, c1 in _groups(Y, g):
            held = hold.view(None, (0, c1 - c0))

            def load(j, Yb, c0=c0, held=held):
                r = panelrange(Y, j)
                backend.convert_copy(held.view(None, (r.start - c0, r.stop - c0)), Yb)
            sweep(load, readpanels(Y), nbuffers=nbuffers, traversal=PanelSteps(list(range(j0, j1))))
            if j0 > 0:
                def update(i, Yb, c0=c0, c1=c1, held=held):
                    r = panelrange(Y, i)
                    backend.rightmul_gemm_scaled(held, Yb, factor.view((r.start, r.stop), (c0, c1)), -1.0, 1.0)
                sweep(update, readpanels(Y), nbuffers=nbuffers, traversal=PanelSteps(list(range(j0))))

            def solve(j, Yb, j0=j0, c0=c0, c1=c1, held=held):
                if j == j0:
                    backend.rdiv_upper(held, factor.view((c0, c1), (c0, c1)), work)
                r = panelrange(Y, j)
                backend.convert_copy(Yb, held.view(None, (r.start - c0, r.stop - c0)))
            sweep(solve, writepanels(Y), nbuffers=nbuffers, traversal=PanelSteps(list(range(j0, j1))))
    finally:
        _plan.checkin_device_buffers(plan, [hold, factor, work])
    return Y


def cholqr2_panelwise(Y, *, nbuffers=None, mixed=False, shift="adaptive", group=None):
    """cholqr2 by column panels (gram_panelwise, host Cholesky,
    rdiv_panelwise), for a basis whose panels do not all fit the host budget.
    Two passes and a third when a Cholesky needed a shift, as cholqr2.
    mixed=True makes the first pass the mixed one of cholqr2_mixed with
    solve="fp64": its Gram on the FP32 matrix cores and its Cholesky in FP32
    (shifted as `shift` says), the solve in FP64. Returns R such that Q R is
    the input; Y holds Q afterwards. `group` is the run length of the panel-wise
    sweeps (panel_group's by default)."""
    def one_pass():
        G = gram_panelwise(Y, nbuffers=nbuffers, group=group)
        R, shifted = cholesky_upper(G, Y.N, Y.k)
        rdiv_panelwise(Y, R, nbuffers=nbuffers, group=group)
        return R, shifted
    if mixed:
        if Y.dtype != np.dtype(np.complex128):
            raise ArgumentError("the mixed-precision pass is for ComplexF64 panel matrices")
        G = gram_panelwise(Y, nbuffers=nbuffers, compute=np.complex64, group=group)
        R32, s = cholesky_upper_fp32(G, Y.N, Y.k, shift)
        R = R32.astype(Y.dtype)
        rdiv_panelwise(Y, R, nbuffers=nbuffers, group=group)
        shifted = s > 0.0
    else:
        R, shifted = one_pass()
    R2, shifted2 = one_pass()
    R = triangular_product(R2, R)
    if shifted or shifted2:
        R3, _ = one_pass()
        R = triangular_product(R3, R)
    return R


def scale(Y, alpha, *, nbuffers=None):
    """Multiplies every entry of Y by alpha in place, in one sweep
    (src/panelblas.jl:303-316). Returns Y."""
    backend = Y.plan.backend
    sweep(lambda _, Yb: backend.scale(Yb, alpha), updatepanels(Y), nbuffers=nbuffers)
    return Y


def axpy(alpha, X, Y, *, nbuffers=None):
    """Y <- alpha X + Y in one sweep (src/panelblas.jl:318-331). Returns Y."""
    shared_plan(X, Y)
    assert_conformal(X, Y)
    if X is Y:
        raise ArgumentError("axpy with X and Y the same matrix is a scaling; call scale(X, 1 + alpha)")
    if not np.issubdtype(Y.dtype, np.complexfloating) and complex(alpha).imag != 0:
        raise ArgumentError(f"cannot add a complex multiple ({alpha}) of a real {Y.dtype} panel matrix")
    backend = Y.plan.backend
    sweep(lambda _, Xb, Yb: backend.axpby(Yb, Xb, alpha, 1.0), readpanels(X), updatepanels(Y), nbuffers=nbuffers)
    return Y


def _reduce(X, Y, nbuffers):
    """Sum over every entry of conj(X) .* Y (|X|^2 when Y is None), one sweep
    over column panels (src/panelblas.jl:333-370). Each step adds its panel's
    sum, formed in FP64 in a fixed order, to a device scalar; the reference's
    N x w accumulator and its host-side sum are not needed."""
    plan = X.plan
    backend = plan.backend
    acc = _plan.checkout_device_buffers(plan, np.complex128, (1, 1), 1)[0]
    work = None
    try:
        work = _plan.checkout_workspace(plan, backend.dot_workspace_bytes(X.N, X.w))
        if Y is None:
            sweep(lambda j, Xb: backend.dot_accumulate(acc, Xb, None, j > 0, work), readpanels(X),
                  nbuffers=nbuffers)
        else:
            sweep(lambda j, Xb, Yb: backend.dot_accumulate(acc, Xb, Yb, j > 0, work), readpanels(X),
                  readpanels(Y), nbuffers=nbuffers)
        if npanels(X) == 0:
            return 0j
        return complex(collect_small(acc, plan)[0, 0])
    finally:
        _plan.checkin_device_buffers(plan, [acc, work])


def dot(X, Y, *, nbuffers=None):
    """sum(conj(X) .* Y) over every entry, in the eltype of X and Y
    (src/panelblas.jl:333-354)."""
    shared_plan(X, Y)
    assert_conformal(X, Y)
    v = _reduce(X, None if X is Y else Y, nbuffers)
    T = X.dtype.type
    return T(v) if np.issubdtype(X.dtype, np.complexfloating) else T(v.real)


def norm(X, p=2, *, nbuffers=None):
    """The Frobenius norm: the plain square root of the sum of |x|^2, without
    the rescaling LinearAlgebra does for arrays (src/panelblas.jl:318-322,
    356-370)."""
    if p != 2:
        raise ArgumentError(f"norm of a PanelMatrix is the Frobenius norm (p = 2) only; p = {p} was asked for")
    R = np.finfo(X.dtype).dtype.type
    return R(np.sqrt(_reduce(X, None, nbuffers).real))


UNIT_ROUNDOFF_FP32 = 2.0 ** -24


def cholesky_upper_fp32(G, N, k, shift="adaptive"):
    """The first pass's preconditioning factor, a Cholesky factorization in
    FP32 of the Gram formed on the FP32 matrix cores. Returns (R, s) with s
    the diagonal shift that was applied.

    shift="fukaya" always factors G + s I with s = 11 (N k + k (k+1)) u32
    ||Y||_2^2 (shifted CholeskyQR, Fukaya et al. 2020; ||Y||_2^2 is the Gram's
    largest eigenvalue). That bound assumes every product of the length-N
    inner products is summed in FP32; at N = 1e7 it exceeds ||Y||_2^2 itself,
    so the factor stops preconditioning and the FP64 pass that follows cannot
    reach FP64 orthogonality. shift="adaptive" (the default) tries the
    unshifted FP32 factorization first and falls back to the Fukaya shift
    only when it fails, as the reference does for its FP64 rule
    (src/panelblas.jl:148-155). This pass has no counterpart in the reference."""
    G = np.asarray(G)
    narrow = np.complex64 if np.iscomplexobj(G) else np.float32
    if shift == "adaptive":
        R = _potrf(G.astype(narrow))
        if R is not None:
            return R, 0.0
    elif shift != "fukaya":
        raise ArgumentError(f"shift must be 'adaptive' or 'fukaya', got {shift!r}")
    lam = scipy.linalg.eigh(G, eigvals_only=True, subset_by_index=[k - 1, k - 1])[0]
    s = 11.0 * (N * k + k * (k + 1)) * UNIT_ROUNDOFF_FP32 * lam
    R = _potrf((G + s * np.eye(k)).astype(narrow))
    if R is None:
        raise ArgumentError(
            f"the {k}×{k} FP32 Gram matrix is not positive definite after the shift {s}; the columns are "
            "numerically dependent")
    return R, s


def cholqr_pass_mixed(Y, nbuffers=None, shift="adaptive", solve="fp32", schedule=None):
    """The mixed first pass: FP32 matrix-core Gram with FP64 sums across row
    chunks and an FP32 Cholesky (shifted when needed), then the triangular
    solve on the FP64 data. solve="fp32" runs the solve's products on the FP32
    matrix cores, which leaves ||Y - Q R|| / ||Y|| near u32 kappa(Y) (Q stays
    orthonormal to FP64 level after the FP64 pass); solve="fp64" solves in
    FP64 with the FP32-derived factor, so Q R reproduces Y to FP64 rounding.
    Returns (R, shift)."""
    if Y.dtype != np.dtype(np.complex128):
        raise ArgumentError("the mixed-precision pass is for ComplexF64 panel matrices")
    if solve not in ("fp64", "fp32"):
        raise ArgumentError(f"solve must be 'fp64' or 'fp32', got {solve!r}")
    if shift not in ("adaptive", "fukaya"):
        raise ArgumentError(f"shift must be 'adaptive' or 'fukaya', got {shift!r}")
    if resolve_schedule(Y, schedule) == "reference":
        G = gram(Y, nbuffers=nbuffers, compute=np.complex64)
        R32, s = cholesky_upper_fp32(G, Y.N, Y.k, shift)
        R = R32.astype(Y.dtype)
        rdiv_rows(Y, R, nbuffers=nbuffers, compute=None if solve == "fp64" else np.complex64, schedule="reference")
        return R, s
    return _device_pass(Y, nbuffers, first=(shift, solve))[:2]


def cholqr2_mixed(Y, *, nbuffers=None, shift="adaptive", solve="fp32", schedule=None):
    """cholqr2 with a mixed-precision first pass (FP32 Gram and Cholesky, the
    solve in the precision `solve` names) as a preconditioner and the second
    pass in FP64. A third FP64 pass runs when either earlier Cholesky needed a
    shift, as shifted CholeskyQR3 prescribes: a shifted first pass leaves Q1
    with a condition number well above 1, which one FP64 pass does not bring
    to FP64 orthogonality. The schedule is cholqr2's: under "reference" every
    pass is a Gram sweep, a host Cholesky and a solve sweep (the first pass's
    Cholesky in FP32); under "fused" the factorizations run on the device and
    the Gram of each next pass rides on the solve sweep off HBM."""
    if Y.dtype != np.dtype(np.complex128):
        raise ArgumentError("the mixed-precision pass is for ComplexF64 panel matrices")
    if solve not in ("fp64", "fp32"):
        raise ArgumentError(f"solve must be 'fp64' or 'fp32', got {solve!r}")
    if shift not in ("adaptive", "fukaya"):
        raise ArgumentError(f"shift must be 'adaptive' or 'fukaya', got {shift!r}")
    if resolve_schedule(Y, schedule) == "reference":
        return _reference_passes(Y, nbuffers, first=(shift, solve))
    return _passes(Y, nbuffers, first=(shift, solve))
