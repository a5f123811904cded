# WITH_HIP for hfsolver: the two-electron part of the closed-shell RHF path on
# the MI355X engine of this repository, selected at build time as WITH_LIBINT
# selects the libint plug-in (src/CMakeLists.txt:67-77). Include it from
# src/CMakeLists.txt once SRC is listed, in place of the gaussians_*.f90
# choice:
#
#     if (WITH_HIP)
#         include(${HFSOLVER_AMD}/hfsolver_amd/fortran/hfx_hip.cmake)
#     elseif (WITH_LIBINT)
#         ...
#
# Reads   SRC           hfsolver's source list, holding scf.f90
#         HFSOLVER_AMD  root of this repository (hfsolver_amd/libhfx.so built)
#         HIP_MODE      "direct" (default) or "store"
# Sets    SRC           with the plug-in's sources (direct mode: scf.f90 replaced)
#         HFX_LIBS      libraries to link
#
# direct: gaussians_direct.f90 (getints2 creates and keeps the engine, and
#         still fills int2 for the MBPT / Green's-function consumers) and
#         scf_direct.f90 (doscf with G(P) from hfx_fock_2e every iteration, no
#         TT tensor). scf.f90 itself is compiled unchanged under the module
#         name hfsolver_scf (a renamed copy in the build tree); scf_direct
#         re-exports its other procedures.
# store:  gaussians_gpu.f90 (getints2 fills int2 with hfx_int2); scf.f90 as is.
if (NOT DEFINED HIP_MODE)
    set(HIP_MODE direct)
endif ()
set(HFX_DIR ${HFSOLVER_AMD}/hfsolver_amd)
if (NOT EXISTS ${HFX_DIR}/libhfx.so)
    message(FATAL_ERROR "WITH_HIP: ${HFX_DIR}/libhfx.so is not built (python -m hfsolver_amd.build)")
endif ()
set(HFX_LIBS ${HFX_DIR}/libhfx.so)
if (HIP_MODE STREQUAL "direct")
    set(_hfx_scf "")
    foreach (_f ${SRC})
        if (_f MATCHES "(^|/)scf\\.f90$")
            set(_hfx_scf ${_f})
        endif ()
    endforeach ()
    if (NOT _hfx_scf)
        message(FATAL_ERROR "WITH_HIP: scf.f90 is not in SRC")
    endif ()
    get_filename_component(_hfx_scf_abs ${_hfx_scf} ABSOLUTE)
    file(READ ${_hfx_scf_abs} _hfx_text)
    string(REGEX REPLACE "(^|\n)module scf\n" "\\1module hfsolver_scf\n" _hfx_text "${_hfx_text}")
    if (NOT _hfx_text MATCHES "module hfsolver_scf")
        message(FATAL_ERROR "WITH_HIP: no 'module scf' line in ${_hfx_scf_abs}")
    endif ()
    file(WRITE ${CMAKE_CURRENT_BINARY_DIR}/hfsolver_scf.f90 "${_hfx_text}")
    set_property(DIRECTORY APPEND PROPERTY CMAKE_CONFIGURE_DEPENDS ${_hfx_scf_abs})
    list(REMOVE_ITEM SRC ${_hfx_scf})
    list(APPEND SRC ${CMAKE_CURRENT_BINARY_DIR}/hfsolver_scf.f90 ${HFX_DIR}/fortran/hfx_gpu.f90
         ${HFX_DIR}/fortran/gaussians_direct.f90 ${HFX_DIR}/fortran/scf_direct.f90)
elseif (HIP_MODE STREQUAL "store")
    list(APPEND SRC ${HFX_DIR}/fortran/gaussians_gpu.f90)
else ()
    message(FATAL_ERROR "WITH_HIP: HIP_MODE must be direct or store, not ${HIP_MODE}")
endif ()
