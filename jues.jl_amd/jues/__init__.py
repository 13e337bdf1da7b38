"""jues -- host-side mirror of JuES.jl's correlation hot path over libjues_hip.so.

Modules mirror the reference's (paths relative to the JuES.jl tree):

* ``Wavefunction``           Backend/Wavefunction.jl   (the ``Wfn`` carrier)
* ``Transformation``         Backend/Transformation.jl (``tei_transform``)
* ``IntegralTransformation`` Backend/IntegralTransformation.jl (``get_eri``)
* ``DF``                     Backend/DF.jl             (``make_bμν`` from the pair ``setup_df`` returns)
* ``MollerPlesset``          MollerPlesset/{RMP2,DF-RMP2}.jl (``do_rmp2``, ``do_df_rmp2``, the MP2 adjoint)
* ``CoupledCluster``         CoupledCluster/{RCCD,mRCCD,RCCSD,mRCCSD,DF-RCCD}.jl

Every numerical call goes through the HIP library; there is no CPU path.
"""
from . import synthetic  # noqa: F401
from ._lib import JuesError, JuesLibraryError, context  # noqa: F401
