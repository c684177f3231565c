"""CPU oracle of the TM5 transport step -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.

A NumPy restatement of pydyn's algorithm for the transport step, written to do the same IEEE
operations in the same order as the reference so that its results are bit-identical to it.
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import it, and only as
the checker: the product (pydyn_amd) never imports, links or runs anything under oracle/.

Modules (each function cites the reference file:line it follows, paths relative to the pydyn
repository root):

    lines       the 1-D slopes-scheme pieces every advection shares (ops.py, kernels.py:64-92)
    advection   dynam0, dynamu (x), dynamv (y), dynamw (z) with their sub-cycle searches
    convection  the convection/diffusion matrix, the Crout inverse and convec
    physics     dkg from Kzz, CH4 injection, CH4 decay, m2phlb1
    cfl         Check_CFL: the mass-only rehearsal and the time-step reduction
    driver      one region's Setup_MassFlow, step (two half steps) and interval
    libm        the fdlibm exp the device also uses (a C restatement, built into _build/)

How each part is pinned (beyond reading): against golden vectors that the reference itself
produced in the build container (tests/golden/), against the reference tests' known answers
(new_valid_timestep, nint), and against invariants (mass conservation, uniform mixing ratio,
the Crout inverse against numpy.linalg.inv, exp against the C library within 1 ulp).
"""
