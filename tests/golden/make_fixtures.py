"""Extract the reference test data this suite pins against into small fixtures.

Reads Bolt.jl's own test data (test/data/*.dat under the reference checkout,
default /root/reference) and writes compact .npz files next to this script:

  recfast_xe.npz      Fortran RECFAST X_e(z) (runtests.jl:39-47), every row
  class_kXXX.npz      CLASS Newtonian-gauge x, Φ, δ_b (runtests.jl:103-145)
  camb_cl.npz         CAMB unlensed D_ℓ TT/TE/EE (runtests.jl:168)
  camb_pk.npz         CAMB linear P(k) at z = 0, k in h/Mpc, P in (Mpc/h)³
                      (test/data/camb_pk_x0.dat, test/camb_v_class_2022.ipynb)

Only data values are stored (inputs and expected outputs).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref="/root/reference"):
    d = os.path.join(ref, "test", "data")
    rec = np.loadtxt(os.path.join(d, "test_recfast_1.dat"), delimiter=",", skiprows=1)
    np.savez_compressed(os.path.join(HERE, "recfast_xe.npz"), z=rec[:, 0], xe=rec[:, 1])
    for tag in ["p01", "p03", "p3", "1p0"]:
        rows = np.loadtxt(os.path.join(d, f"zack_N_class_px_k{tag}_nofluid_nonu.dat"))
        kclass = rows[1][0]
        rows = np.delete(rows, 1, axis=0)
        # CLASS keys: x, d_g, d_b, d_cdm, d_ur, d_ncdm, d_tot, phi, psi, ...
        np.savez_compressed(os.path.join(HERE, f"class_k{tag}.npz"), k_hMpc=kclass, x=rows[0][::-1],
                            delta_b=rows[2][::-1], phi=rows[6][::-1])
    camb = np.loadtxt(os.path.join(d, "camb_rough_ttteee_unlensed.dat"))
    np.savez_compressed(os.path.join(HERE, "camb_cl.npz"), ell=camb[0], tt=camb[1], te=camb[2], ee=camb[3])
    pk = np.loadtxt(os.path.join(d, "camb_pk_x0.dat"))
    np.savez_compressed(os.path.join(HERE, "camb_pk.npz"), k_h=pk[0], P_h=pk[1])


if __name__ == "__main__":
    main(*sys.argv[1:])
