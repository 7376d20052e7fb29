"""The oracle: a CPU restatement of Funicular.jl's reference CPU backend for
the cholqr2! / gram / rdiv_rows! / project / panelmul! / GhostPanels / sweep!
path, used only to check the HIP backend.

TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import, call or link anything here, and only as the
checker or the reported CPU baseline -- never as the thing measured or
shipped. The product (funicular.jl_amd/) never imports it.

Contents, each function citing the reference file:line it follows:
  plan.py       integer schedule arithmetic (src/plan.jl, src/executor.jl)
  panelblas.py  gram / cholesky_upper / rdiv / cholqr2 / rightmul / project /
                panelmul in the reference's step order, FP64 numpy + LAPACK
  operators.py  the configuration operators applied densely on the CPU
  ghost.c       the ghost-panel stream (counter-based SplitMix64 + 256-layer
                ziggurat with fixed-sequence exp/log), built by build.py
  ghost.py      ctypes wrapper of ghost.c
  adjoint.py    the adjoints of cholqr2 and project (dense FP64)
  build.py      compiles ghost.c into _build/

Pinning (what makes each function trustworthy; see tests/test_oracle_*.py):
  plan.py       the reference's literal known answers (test/test_plan.jl:18-38,
                :68-88, test/test_executor.jl:287-301) and its properties
                (test/test_plan.jl:40-66) over seeded random cases.
  panelblas.py  closed forms and textbook LAPACK routines: Gram vs a brute-force
                mpmath product on tiny inputs, TRSM residuals, Cholesky shift
                known answers (test/test_panelblas.jl:327-345), QR invariants
                (Q^H Q = I, Q R = Y, R upper with positive diagonal), project
                vs Q^H (G Q) formed explicitly, nbuffers-independent schedule.
  operators.py  stencils vs an explicit sparse/dense matrix built index by
                index, adjoint identity <y, Gx> = <G^H y, x>.
  ghost.c       SplitMix64 published outputs; det_exp/det_log within 2 ulp of
                libm over a sweep; the ziggurat layer-area identity; moments
                and a Kolmogorov-Smirnov test of the normals.
                The reference's own ghost bits (Julia hash + Xoshiro256++ +
                randn!, src/ghosts.jl:58,77,100) cannot be produced here:
                bit parity with the Julia stream is UNPINNED (no Julia in
                this image); the stream is this build's own, specified in
                ghost.c, and pinned by the checks above.
"""
