# xss_amdConfig.cmake -- CMake package of the MI355X sort/select library, the
# counterpart of the reference's pkg-config file x86simdsortcpp.pc and its
# meson dependency x86simdsortcpp_dep (meson.build:72-83).
#
#   find_package(xss_amd CONFIG REQUIRED)      # -Dxss_amd_DIR=<repo>/x86-simd-sort_amd/cmake
#   target_link_libraries(app PRIVATE xss_amd::x86simdsortcpp)
#
# The target carries include/ (x86simdsort.h, x86simdsort-static-incl.h,
# xss_gpu.h) and libxss_amd.so, which exports the reference's
# x86simdsort:: specialisations and the C ABI.  Build the library first:
# python x86-simd-sort_amd/build.py
get_filename_component(_xss_amd_pkg "${CMAKE_CURRENT_LIST_DIR}/.." ABSOLUTE)
get_filename_component(_xss_amd_root "${_xss_amd_pkg}/.." ABSOLUTE)
set(_xss_amd_lib "${_xss_amd_pkg}/libxss_amd.so")

if(NOT EXISTS "${_xss_amd_lib}")
  set(xss_amd_FOUND FALSE)
  set(xss_amd_NOT_FOUND_MESSAGE
      "${_xss_amd_lib} is not built; run: python ${_xss_amd_pkg}/build.py")
  return()
endif()

if(NOT TARGET xss_amd::x86simdsortcpp)
  add_library(xss_amd::x86simdsortcpp SHARED IMPORTED)
  set_target_properties(xss_amd::x86simdsortcpp PROPERTIES
    IMPORTED_LOCATION "${_xss_amd_lib}"
    IMPORTED_SONAME "libxss_amd.so"
    INTERFACE_INCLUDE_DIRECTORIES "${_xss_amd_root}/include"
    INTERFACE_COMPILE_FEATURES cxx_std_17)
endif()

set(xss_amd_INCLUDE_DIRS "${_xss_amd_root}/include")
set(xss_amd_LIBRARIES xss_amd::x86simdsortcpp)
