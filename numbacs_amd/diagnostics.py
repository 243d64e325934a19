"""FTLE and Cauchy-Green diagnostics (mirrors numbacs.diagnostics).

Reference: ftle_grid_2D (diagnostics.py:21-65), C_tensor_2D
(diagnostics.py:67-110), C_eig_aux_2D (diagnostics.py:113-197), C_eig_2D
(diagnostics.py:200-244) and ftle_from_eig (diagnostics.py:247-269).
"""
import numpy as np

from . import _lib


def _fm_dev(flowmap):
    t = _lib.torch()
    if isinstance(flowmap, t.Tensor):
        fm = _lib.to_dev(flowmap)
    else:
        fm = _lib.to_dev(np.asarray(flowmap))
    if fm.ndim != 3 or fm.shape[2] != 2:
        raise ValueError("flowmap must have shape (nx, ny, 2)")
    return fm


def _mask_dev(mask, nx, ny):
    if mask is None:
        return None
    m = np.asarray(mask)
    if m.shape != (nx, ny):
        raise ValueError(f"mask must have shape {(nx, ny)}")
    return _lib.to_dev(m.astype(np.uint8), np.uint8)


def ftle_grid_2d_device(fmd, T, dx, dy, maskd=None, out=None):
    nx, ny = int(fmd.shape[0]), int(fmd.shape[1])
    if out is None:
        out = _lib.empty((nx, ny))
    _lib.check(_lib.load().nbx_ftle_grid_2d(_lib.ptr(fmd), nx, ny, float(T), float(dx), float(dy),
                                            _lib.ptr(maskd), _lib.ptr(out), _lib.stream_ptr()))
    return out


def ftle_eig_2d_device(fmd, T, dx, dy, maskd=None, ftle=None, evec=None):
    """Fused C_eig_2D + ftle_from_eig: (ftle (nx,ny), eigvec_max (nx,ny,2))."""
    nx, ny = int(fmd.shape[0]), int(fmd.shape[1])
    if ftle is None:
        ftle = _lib.empty((nx, ny))
    if evec is None:
        evec = _lib.empty((nx, ny, 2))
    _lib.check(_lib.load().nbx_ftle_eig_2d(_lib.ptr(fmd), nx, ny, float(T), float(dx), float(dy),
                                           _lib.ptr(maskd), _lib.ptr(ftle), _lib.ptr(evec),
                                           _lib.stream_ptr()))
    return ftle, evec


def ftle_grid_2D(flowmap, T, dx, dy, mask=None):
    """FTLE field (nx, ny) of a flow map (nx, ny, 2) for integration time T:
    central differences on the interior, 0 on the border and at masked
    points, log(lambda_max)/(2|T|) where lambda_max > 1 and 0 elsewhere."""
    fm = _fm_dev(flowmap)
    nx, ny = int(fm.shape[0]), int(fm.shape[1])
    if nx * ny == 0:
        return np.zeros((nx, ny))
    maskd = _mask_dev(mask, nx, ny)
    out = ftle_grid_2d_device(fm, T, dx, dy, maskd)
    return _lib.to_host(out)


def C_eig_2D(flowmap, dx, dy, mask=None):
    """Eigenvalues (nx, ny, 2), ascending, and eigenvectors (nx, ny, 2, 2)
    (columns, eigvecs[..., :, 1] for the largest) of the Cauchy-Green
    tensor.  Eigenvector signs may differ from LAPACK's; the products the
    reference tests compare (V diag(l) V^T) and the ridge points are
    sign-invariant."""
    fm = _fm_dev(flowmap)
    nx, ny = int(fm.shape[0]), int(fm.shape[1])
    if nx * ny == 0:
        return np.zeros((nx, ny, 2)), np.zeros((nx, ny, 2, 2))
    ev = _lib.empty((nx, ny, 2))
    V = _lib.empty((nx, ny, 2, 2))
    maskd = _mask_dev(mask, nx, ny)  # held until the kernel has been enqueued and copied back
    _lib.check(_lib.load().nbx_C_eig_2d(_lib.ptr(fm), nx, ny, float(dx), float(dy),
                                        _lib.ptr(maskd), _lib.ptr(ev), _lib.ptr(V),
                                        _lib.stream_ptr()))
    return _lib.to_host(ev), _lib.to_host(V)


def _fm_aux_dev(flowmap_aux):
    t = _lib.torch()
    fa = _lib.to_dev(flowmap_aux if isinstance(flowmap_aux, t.Tensor) else
                     np.asarray(flowmap_aux))
    if fa.ndim != 4 or fa.shape[3] != 2 or fa.shape[2] not in (4, 5):
        raise ValueError("flowmap_aux must have shape (nx, ny, n_aux, 2) with n_aux 4 or 5")
    return fa


def C_tensor_2D(flowmap_aux, dx, dy, h=1e-5, mask=None):
    """(C11, C12, C22) of the Cauchy-Green tensor from the aux-grid flow map,
    (nx, ny, 3); zero within two points of the border and at masked points
    (dx, dy are unused, as in the reference)."""
    fa = _fm_aux_dev(flowmap_aux)
    nx, ny, naux = int(fa.shape[0]), int(fa.shape[1]), int(fa.shape[2])
    if nx * ny == 0:
        return np.zeros((nx, ny, 3))
    C = _lib.empty((nx, ny, 3))
    maskd = _mask_dev(mask, nx, ny)
    _lib.check(_lib.load().nbx_C_tensor_aux_2d(_lib.ptr(fa), nx, ny, naux, float(h),
                                               _lib.ptr(maskd), _lib.ptr(C), _lib.stream_ptr()))
    return _lib.to_host(C)


def C_eig_aux_2D(flowmap_aux, dx, dy, h=1e-5, eig_main=True, mask=None):
    """Eigenvalues (nx, ny, 2) and eigenvectors (nx, ny, 2, 2) from the
    aux-grid flow map: with eig_main the eigenvectors come from the aux-grid
    tensor and the eigenvalues from the main-grid tensor of the centre
    particles (zero within two points of the border); without it both come
    from the aux-grid tensor (zero on the border).  Eigenvector signs may
    differ from LAPACK's, as for C_eig_2D."""
    fa = _fm_aux_dev(flowmap_aux)
    nx, ny, naux = int(fa.shape[0]), int(fa.shape[1]), int(fa.shape[2])
    if nx * ny == 0:
        return np.zeros((nx, ny, 2)), np.zeros((nx, ny, 2, 2))
    if eig_main and naux != 5:
        raise ValueError("eig_main needs the centre particle: flowmap_aux with n_aux = 5")
    ev = _lib.empty((nx, ny, 2))
    V = _lib.empty((nx, ny, 2, 2))
    maskd = _mask_dev(mask, nx, ny)
    _lib.check(_lib.load().nbx_C_eig_aux_2d(_lib.ptr(fa), nx, ny, naux, float(dx), float(dy),
                                            float(h), int(bool(eig_main)), _lib.ptr(maskd),
                                            _lib.ptr(ev), _lib.ptr(V), _lib.stream_ptr()))
    return _lib.to_host(ev), _lib.to_host(V)


def ftle_from_eig(eigval_max, T):
    """FTLE from the largest Cauchy-Green eigenvalue (diagnostics.py:247-269):
    log(lambda)/(2|T|) where lambda > 1, else 0.  Elementwise host-side
    array arithmetic, as in the reference (which is plain numpy here too)."""
    eigval_max = np.asarray(eigval_max, dtype=np.float64)
    ftle = np.zeros(eigval_max.shape)
    sel = np.nonzero(eigval_max > 1)
    ftle[sel] = np.log(eigval_max[sel]) / (2 * abs(T))
    return ftle
