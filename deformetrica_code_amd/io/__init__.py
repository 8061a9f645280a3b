"""File formats of the reference's callers of the kernel-sum engine.

``vtk``    ASCII legacy VTK POLYDATA meshes and polylines (the reference reads them with
           vtkPolyDataReader, DeformableObjectReader.cxx:58-145)
``dlm``    whitespace-delimited matrix text (src/io/MatrixDLM.cxx)
``state``  the optimizer snapshot, a Boost text archive
           (src/support/utilities/SerializeDeformationState.h)
"""
from .dlm import (read_matrix_dlm, read_multiple_matrix_dlm, write_matrix_dlm,  # noqa: F401
                  write_multiple_matrix_dlm)
from .state import (DeformationState, pop_fast_gradient_ascent, pop_gradient_ascent,  # noqa: F401
                    push_fast_gradient_ascent, push_gradient_ascent, states_close)
from .vtk import PolyData, read_vtk, write_vtk  # noqa: F401
