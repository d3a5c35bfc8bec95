"""CPU oracle for the two-phase D2Q9 lattice-Boltzmann step (TEST INFRASTRUCTURE ONLY).

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``fingering_dynamics_amd``) never imports,
calls or links anything in here and fails loudly when its HIP library is
missing.

Contents
--------
``lbm64``      FP64 numpy restatement of the per-step update of
               ``lattice_boltzmann/fingering_periodic.py`` (and of the walled
               sibling ``lattice_boltzmann/fingering.py``), op for op: boolean
               mask gathers/scatters, ``np.roll`` streaming, deep copies of the
               post-collision populations, per-direction half-way bounce-back
               masks and the Zou-He column updates.
``geometry64`` restatement of the obstacle front-end of
               ``lattice_boltzmann/create_block.py`` and of the bounce-back mask
               unions of ``lattice_boltzmann/bounce_back.py``.
``params64``   the physical -> lattice unit conversion of
               ``fingering_periodic.py:15-44`` / ``fingering.py:16-50``.

How each function is pinned (closed forms, invariants, brute force) is listed
in the header of each module.  One part is *parity unpinned*: the circle
rasteriser used by ``geometry64.circle_outline`` stands in for
``cv2.circle`` (create_block.py:19-21), which is not installed here and has no
golden vectors in the reference; the classifier run on top of it is pinned by
invariants only.
"""
