"""MI355X-native implementation of the self-absorption-corrected XRF
reconstruction hot path of gmysage/FL_correction: attenuation maps, the rotated
emission projector and its transpose, MLEM/OSL-Huber, the adjoint of the
attenuated forward model, and batched execution -- hand-written HIP kernels for
gfx950 behind the reference's Python API."""
__version__ = "0.1.0"
