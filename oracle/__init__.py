"""FP64 NumPy oracle for the dense smoke/fire step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``fluxfx_amd``) may import,
call or execute anything here.  The only permitted users are ``tests/``,
``__graft_entry__.smoke()`` (as the checker of one small GPU run) and the
``cpu_baseline`` leg of ``bench.py`` (timed as the CPU reference).

Every function restates one piece of the reference algorithm from its source
text and cites the reference file:line it follows (paths relative to the
reference repository root).  Arrays are C-ordered ``[z, y, x]`` so that a
flat view is the reference's x-fast index ``x + nx*(y + ny*z)``
(physics/reference.py:3); face arrays are U ``[nz, ny, nx+1]``,
V ``[nz, ny+1, nx]``, W ``[nz+1, ny, nx]`` (physics/config.py:21-23).

Parity pinning (how each function is checked against something other than a
retyping of itself) is listed in ``tests/test_oracle_*.py``.  Functions with
no reference counterpart (tolerance-driven multigrid, mixed precision,
U-Net) are pinned by closed forms, brute force and finite differences.
"""
