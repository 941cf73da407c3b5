"""Parity oracle for the split-step BPM and Debye-Wolf hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from this package, and
only as the checker (or as the timed CPU baseline), never as the thing that is
measured or shipped.  The product package ``bpm_amd`` never imports it.

Modules
-------
bpm_oracle   numpy restatement of ``bpm/bpm/bpm_3d.py`` (``_bpm_3d``,
             ``_bpm_3d_split``), the OpenCL kernels in
             ``bpm/bpm/kernels/bpm_3d_kernels.cl`` and the class path
             ``bpm/bpmclass/_bpm3d_ocl.py``; complex64 ("c64", stand-in for
             the reference float32 OpenCL path) and complex128 ("f64").
psf_oracle   numpy/scipy restatement of ``bpm/psf/_focus_fields_debye.py`` and
             ``bpm/psf/kernels/psf_debye.cl`` (all four kernel variants and
             the octant mirror); float32 and float64.
"""
