"""CPU oracle for the FIT-TD stepping path — TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference algorithm of the path that
``magnelio_amd`` runs on the GPU.  It exists only to check the GPU results:
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, nothing in ``magnelio_amd`` does, and no product
code path falls back to it.

Each function cites the reference file and line range it follows (paths are
relative to the reference repository root, ``src/`` = ``src/magnelio/``).

Modules
-------
yee       field shapes, the fused curl updates (array-stencil and Numba
          operation orders), the stored-energy sum
mesh      grid lines, the AABB material fill, PEC edge mask, mass matrices,
          alpha/beta update coefficients, port-plane flattening
cpml      CFS-PML profiles and the psi recursion
dtbc      DTBC symbol, kernel and boundary recursion
ports     port-plane maps, discrete rectangular-waveguide modes, the modal
          port step (projection, DTBC, write-back) and the V/I recorder
sparams   port-signal DFT, power waves and S-parameters
monitors  running DFT of field samples
lumped    lumped RLC element step and trapezoidal companions
tfsf      plane-wave TF/SF patches
dispersion  pole-residue ADE update
fdtd      the leapfrog step order tying the above together

Parity status of each function (pinned by which check) is listed in the
module docstrings; anything not pinned says "parity unpinned".
"""
