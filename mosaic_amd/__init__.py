"""mosaic_amd: the MOSAIC kinematic-scattering and propagation hot path on MI355X (gfx950).

Host objects mirror the reference's interface -- ``beam.beam``,
``detector.detector``, ``stage.stage``, ``optics.optics``, ``sample.sample``
(the reference's ``Beam.beam``, ``Detector.detector`` ...); the arithmetic
runs in hand-written HIP kernels in ``libmosaic_amd.so`` behind a C ABI
(include/mosaic_amd.h).
"""
