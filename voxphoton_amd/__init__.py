"""MI355X-native (gfx950) FFT-VIE solve path of VoxPhoton.

The N-operator matvec, the circulant preconditioners and the GMRES Krylov
loop run in libvoxphoton.so (hand-written HIP); this package is the host-side
mirror of the reference's MATLAB call surface.  Importing a submodule that
needs the library raises ImportError when it has not been built.
"""
__all__ = ['vie', 'configs', 'mor', 'krylov_dr']
