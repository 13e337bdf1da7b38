"""jues -- host-side mirror of JuES.jl's correlation hot path over libjues_hip.so.

Modules mirror the reference's (paths relative to the JuES.jl tree):

* ``Wavefunction``           Backend/Wavefunction.jl   (the ``Wfn`` carrier)
* ``Transformation``         Backend/Transformation.jl (``tei_transform``)
* ``IntegralTransformation`` Backend/IntegralTransformation.jl (``get_eri``)
* ``MollerPlesset``          MollerPlesset/RMP2.jl     (``do_rmp2`` + the MP2 adjoint)
* ``CoupledCluster``         CoupledCluster/{RCCD,mRCCD,RCCSD,mRCCSD}.jl

Every numerical call goes through the HIP library; there is no CPU path.
"""
from . import synthetic  # noqa: F401
from ._lib import JuesError, JuesLibraryError, context  # noqa: F401
