"""Argument layout of every LSM3D_* entry point the library exports.

LSMLIB's toolbox passes every argument by reference (Fortran-77 calling
convention, C prototypes in src/toolbox/*/lsm_*3d.h).  Each entry below spells
one prototype as a sequence of argument kinds:

  A  LSMLIB_REAL*      array (column-major, i fastest)
  B  6 x const int*    ghost/fill/include box  ilo, ihi, jlo, jhi, klo, khi
  R  const LSMLIB_REAL* scalar input
  O  LSMLIB_REAL*      scalar output
  I  const int*        integer input
  X  int*              index list (index_x/y/z, index_outer): device memory
                       for the product library
  N  int*              per-level list limits n_lo / n_hi: host memory
  U  unsigned char*    narrow-band mark array on its box (compared as the
                       Fortran integer*1, i.e. signed)
  M  const unsigned char*  one mark value
  K  int*              integer output

The same table drives the ctypes marshalling for the product library and lets
tests check that every declared symbol is exported.
"""

# (symbol, kinds, reference prototype file:line)
_TABLE = [
    # spatial derivatives -- lsm_spatial_derivatives3d.h
    ("lsm3dhjeno1_", "AAAB AAAB AB AB B RRR", "lsm_spatial_derivatives3d.h:83"),
    ("lsm3dhjeno2_", "AAAB AAAB AB AB AB B RRR", "lsm_spatial_derivatives3d.h:159"),
    ("lsm3dhjeno3_", "AAAB AAAB AB AB AB AB B RRR", "lsm_spatial_derivatives3d.h:243"),
    ("lsm3dhjweno5_", "AAAB AAAB AB AB B RRR", "lsm_spatial_derivatives3d.h:332"),
    ("lsm3dupwindhjeno1_", "AAAB AB AAAB AB B RRR", "lsm_spatial_derivatives3d.h:395"),
    ("lsm3dupwindhjeno2_", "AAAB AB AAAB AB AB B RRR", "lsm_spatial_derivatives3d.h:459"),
    ("lsm3dupwindhjeno3_", "AAAB AB AAAB AB AB AB B RRR", "lsm_spatial_derivatives3d.h:531"),
    ("lsm3dupwindhjweno5_", "AAAB AB AAAB AB B RRR", "lsm_spatial_derivatives3d.h:608"),
    ("lsm3dcentralgradorder2_", "AAAB AB B RRR", "lsm_spatial_derivatives3d.h:673"),
    ("lsm3dcentralgradorder4_", "AAAB AB B RRR", "lsm_spatial_derivatives3d.h:723"),
    ("lsm3dlaplacianorder2_", "AB AB B RRR", "lsm_spatial_derivatives3d.h:777"),
    ("lsm3dgradientmagnitude_", "AAAAB B", "lsm_spatial_derivatives3d.h:939"),
    ("lsm3dphiupwindgradf_", "AAAB AAAB AAAB AB B", "lsm_spatial_derivatives3d.h:831"),
    ("lsm3daveragegradphi_", "AAAB AAAB AAAB B", "lsm_spatial_derivatives3d.h:893"),
    # level set equation RHS -- lsm_level_set_evolution3d.h
    ("lsm3dzerooutlevelseteqnrhs_", "AB", "lsm_level_set_evolution3d.h:64"),
    ("lsm3daddadvectiontermtolserhs_", "AB AAAB AAAB B", "lsm_level_set_evolution3d.h:94"),
    ("lsm3daddnormalveltermtolserhs_", "AB AAAB AAAB AB B", "lsm_level_set_evolution3d.h:151"),
    ("lsm3daddconstnormalveltermtolserhs_", "AB AAAB AAAB R B", "lsm_level_set_evolution3d.h:217"),
    ("lsm3daddexternalandnormalveltermtolserhs_", "AB AAAB AAAB AAAAB B",
     "lsm_level_set_evolution3d.h:398"),
    ("lsm3daddconstcurvtermtolserhs_", "AB AAAB AAAAAAB R B", "lsm_level_set_evolution3d.h:277"),
    # TVD Runge-Kutta -- lsm_tvd_runge_kutta3d.h
    ("lsm3drk1step_", "AB AB AB B R", "lsm_tvd_runge_kutta3d.h:59"),
    ("lsm3dtvdrk2stage1_", "AB AB AB B R", "lsm_tvd_runge_kutta3d.h:108"),
    ("lsm3dtvdrk2stage2_", "AB AB AB AB B R", "lsm_tvd_runge_kutta3d.h:154"),
    ("lsm3dtvdrk3stage1_", "AB AB AB B R", "lsm_tvd_runge_kutta3d.h:210"),
    ("lsm3dtvdrk3stage2_", "AB AB AB AB B R", "lsm_tvd_runge_kutta3d.h:256"),
    ("lsm3dtvdrk3stage3_", "AB AB AB AB B R", "lsm_tvd_runge_kutta3d.h:309"),
    # reinitialization -- lsm_reinitialization3d.h
    ("lsm3dcomputereinitializationeqnrhs_", "AB AB AB AAAB AAAB B RRR I",
     "lsm_reinitialization3d.h:77"),
    ("lsm3dcomputeorthogonalizationeqnrhs_", "AB AAAB AAAB AB AAAB AAAB B RRR",
     "lsm_reinitialization3d.h:163"),
    ("lsm3dcomputedistanceforsubcellfix_", "AAB B RRR", "lsm_reinitialization3d.h:247"),
    ("lsm3dcomputereinitializationeqnrhssubcellfixorder1_", "AB AB AAB AAAB AAAB B RRR",
     "lsm_reinitialization3d.h:295"),
    # normals -- lsm_geometry3d.h; field extension -- lsm_field_extension3d.h
    ("lsm3dcomputeunitnormal_", "AAAB AAAB B", "lsm_geometry3d.h:74"),
    ("lsm3dcomputesignedunitnormal_", "AAAB AAAB AB B RRR", "lsm_geometry3d.h:127"),
    ("lsm3dcomputefieldextensioneqnrhs_", "AB AB AB AAAB AAAB B RRR", "lsm_field_extension3d.h:65"),
    # reductions -- lsm_utilities3d.h, lsm_geometry3d.h
    ("lsm3dmaxnormdiff_", "O AB AB B", "lsm_utilities3d.h:87"),
    ("lsm3dmaxnormdiffcontrolvolume_", "O AB AB AB I B", "lsm_utilities3d.h:533"),
    ("lsm3dcomputestableadvectiondt_", "O AAAB B RRR R", "lsm_utilities3d.h:221"),
    ("lsm3dcomputestableadvectiondtcontrolvolume_", "O AAAB AB I B RRR R",
     "lsm_utilities3d.h:585"),
    ("lsm3dcomputestablenormalveldt_", "O AB AAAB AAAB B RRR R", "lsm_utilities3d.h:268"),
    ("lsm3dcomputestableconstnormalveldt_", "O R AAAB AAAB B RRR R", "lsm_utilities3d.h:331"),
    ("lsm3dvolumeintegralphilessthanzero_", "O AB AB B RRR R", "lsm_utilities3d.h:383"),
    ("lsm3dvolumeintegralphigreaterthanzero_", "O AB AB B RRR R", "lsm_utilities3d.h:430"),
    ("lsm3dvolumeregionphilessthanzero_", "O AB B RRR R", "lsm_geometry3d.h:180"),
    ("lsm3dvolumeregionphigreaterthanzero_", "O AB B RRR R", "lsm_geometry3d.h:217"),
    # boundary conditions -- lsm_boundary_conditions3d.h
    ("lsm3dlinearextrapolation_", "AB B I", "lsm_boundary_conditions3d.h:71"),
    ("lsm3dsignedlinearextrapolation_", "AB B I", "lsm_boundary_conditions3d.h:111"),
    ("lsm3dcopyextrapolation_", "AB B I", "lsm_boundary_conditions3d.h:135"),
    ("lsm3dhomogeneousneumanneno1_", "AB B I", "lsm_boundary_conditions3d.h:161"),
    ("lsm3dhomogeneousneumanneno2_", "AB B I", "lsm_boundary_conditions3d.h:186"),
    ("lsm3dhomogeneousneumanneno3_", "AB B I", "lsm_boundary_conditions3d.h:211"),
    ("lsm3dhomogeneousneumannweno5_", "AB B I", "lsm_boundary_conditions3d.h:236"),
    # narrow band -- lsm_localization3d.h
    ("lsm3ddeterminenarrowband_", "AB UB XXX II NN X II KKKK RR I", "lsm_localization3d.h:89"),
    ("lsm3dmarknarrowbandboundarylayer_", "UB B M", "lsm_localization3d.h:138"),
    ("lsm3ddeterminenarrowbandfrommask_", "AB UB XXX II NN I I", "lsm_localization3d.h:187"),
    ("lsm3dmultiplycutofflserhslocal_", "AAB XXX II UB M RR", "lsm_localization3d.h:230"),
    ("lsm3dcheckouternarrowbandlayer_", "K AB XXX II X II II II", "lsm_localization3d.h:270"),
    ("lsm3dimposemasklocal_", "AB AB AB XXX II", "lsm_localization3d.h:307"),
    ("lsm3dcopydatalocal_", "AB AB XXX II", "lsm_localization3d.h:347"),
    # narrow-band kernels -- lsm_*3d_local.h
    ("lsm3dhjeno1local_", "AAAB AAAB AB AB RRR XXX IIII UB MM", "lsm_spatial_derivatives3d_local.h:78"),
    ("lsm3dhjeno2local_", "AAAB AAAB AB AB AB RRR XXX IIIIII UB MMM",
     "lsm_spatial_derivatives3d_local.h:159"),
    ("lsm3dcentralgradorder2local_", "AAAB AB RRR XXX II UB M", "lsm_spatial_derivatives3d_local.h:240"),
    ("lsm3dcentralgradorder4local_", "AAAB AB RRR XXX II UB M", "lsm_spatial_derivatives3d_local.h:293"),
    ("lsm3dlaplacianorder2local_", "AB AB RRR XXX II UB M", "lsm_spatial_derivatives3d_local.h:345"),
    ("lsm3dcomputeavegradphilocal_", "O AB RRR XXX II UB M", "lsm_spatial_derivatives3d_local.h:397"),
    ("lsm3dgradientmagnitudelocal_", "AAAAB XXX II UB M", "lsm_spatial_derivatives3d_local.h:438"),
    ("lsm3drk1steplocal_", "AB AB AB R XXX II UB M", "lsm_tvd_runge_kutta3d_local.h:35"),
    ("lsm3dtvdrk2stage1local_", "AB AB AB R XXX II UB M", "lsm_tvd_runge_kutta3d_local.h:73"),
    ("lsm3dtvdrk2stage2local_", "AB AB AB AB R XXX II UB M", "lsm_tvd_runge_kutta3d_local.h:111"),
    ("lsm3dtvdrk3stage1local_", "AB AB AB R XXX II UB M", "lsm_tvd_runge_kutta3d_local.h:155"),
    ("lsm3dtvdrk3stage2local_", "AB AB AB AB R XXX II UB M", "lsm_tvd_runge_kutta3d_local.h:192"),
    ("lsm3dtvdrk3stage3local_", "AB AB AB AB R XXX II UB M", "lsm_tvd_runge_kutta3d_local.h:237"),
    ("lsm3dzerooutlevelseteqnrhslocal_", "AB XXX II", "lsm_level_set_evolution3d_local.h:67"),
    ("lsm3daddadvectiontermtolserhslocal_", "AB AAAB AAAB XXX II UB M",
     "lsm_level_set_evolution3d_local.h:102"),
    ("lsm3daddnormalveltermtolserhslocal_", "AB AAAB AAAB AB XXX II UB M",
     "lsm_level_set_evolution3d_local.h:166"),
    ("lsm3daddconstnormalveltermtolserhslocal_", "AB AAAB AAAB R XXX II UB M",
     "lsm_level_set_evolution3d_local.h:237"),
    ("lsm3daddconstcurvtermtolserhslocal_", "AB AAAB AAAAAAB R XXX II UB M",
     "lsm_level_set_evolution3d_local.h:300"),
    ("lsm3daddconstprecomputedcurvtermtolserhslocal_", "AB AB AB R XXX II UB M",
     "lsm_level_set_evolution3d_local.h:374"),
    ("lsm3daddexternalandnormalveltermtolserhslocal_", "AB AAAB AAAB AAAAB XXX II UB M",
     "lsm_level_set_evolution3d_local.h:449"),
    ("lsm3dcomputereinitializationeqnrhslocal_", "AB AB AB AAAB AAAB RRR I XXX II UB M",
     "lsm_reinitialization3d_local.h:41"),
    ("lsm3dcomputeorthogonalizationeqnrhslocal_", "AB AAAB AAAB AB AAAB AAAB RRR XXX II UB M",
     "lsm_reinitialization3d_local.h:99"),
    ("lsm3dmaxnormdifflocal_", "O AB AB XXX II UB M", "lsm_utilities3d_local.h:65"),
    ("lsm3dcomputestableadvectiondtlocal_", "O AAAB RRR R XXX II UB M", "lsm_utilities3d_local.h:112"),
    ("lsm3dcomputestablenormalveldtlocal_", "O AB AAAB AAAB RRR R XXX II UB M",
     "lsm_utilities3d_local.h:169"),
    ("lsm3dcomputestablenormalveldtcontrolvolumelocal_", "O AB AAAB AAAB AB I RRR R XXX II UB M",
     "lsm_utilities3d_local.h:241"),
    ("lsm3dcomputestableconstnormalveldtlocal_", "O R AAAB AAAB RRR R XXX II UB M",
     "lsm_utilities3d_local.h:322"),
]

SIGNATURES = {name: kinds.replace(" ", "") for name, kinds, _ in _TABLE}
CITATIONS = {name: cite for name, _, cite in _TABLE}

# The C macro name each F77 symbol is reached through in the reference headers.
MACRO_OF = {
    "lsm3dhjeno1_": "LSM3D_HJ_ENO1",
    "lsm3dhjeno2_": "LSM3D_HJ_ENO2",
    "lsm3dhjeno3_": "LSM3D_HJ_ENO3",
    "lsm3dhjweno5_": "LSM3D_HJ_WENO5",
    "lsm3dupwindhjeno1_": "LSM3D_UPWIND_HJ_ENO1",
    "lsm3dupwindhjeno2_": "LSM3D_UPWIND_HJ_ENO2",
    "lsm3dupwindhjeno3_": "LSM3D_UPWIND_HJ_ENO3",
    "lsm3dupwindhjweno5_": "LSM3D_UPWIND_HJ_WENO5",
    "lsm3dcentralgradorder2_": "LSM3D_CENTRAL_GRAD_ORDER2",
    "lsm3dcentralgradorder4_": "LSM3D_CENTRAL_GRAD_ORDER4",
    "lsm3dlaplacianorder2_": "LSM3D_LAPLACIAN_ORDER2",
    "lsm3dgradientmagnitude_": "LSM3D_GRADIENT_MAGNITUDE",
    "lsm3dphiupwindgradf_": "LSM3D_PHI_UPWIND_GRAD_F",
    "lsm3daveragegradphi_": "LSM3D_AVERAGE_GRAD_PHI",
    "lsm3dcomputeunitnormal_": "LSM3D_COMPUTE_UNIT_NORMAL",
    "lsm3dcomputesignedunitnormal_": "LSM3D_COMPUTE_SIGNED_UNIT_NORMAL",
    "lsm3dcomputefieldextensioneqnrhs_": "LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS",
    "lsm3dzerooutlevelseteqnrhs_": "LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS",
    "lsm3daddadvectiontermtolserhs_": "LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS",
    "lsm3daddnormalveltermtolserhs_": "LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS",
    "lsm3daddconstnormalveltermtolserhs_": "LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS",
    "lsm3daddexternalandnormalveltermtolserhs_":
        "LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS",
    "lsm3daddconstcurvtermtolserhs_": "LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS",
    "lsm3drk1step_": "LSM3D_RK1_STEP",
    "lsm3dtvdrk2stage1_": "LSM3D_TVD_RK2_STAGE1",
    "lsm3dtvdrk2stage2_": "LSM3D_TVD_RK2_STAGE2",
    "lsm3dtvdrk3stage1_": "LSM3D_TVD_RK3_STAGE1",
    "lsm3dtvdrk3stage2_": "LSM3D_TVD_RK3_STAGE2",
    "lsm3dtvdrk3stage3_": "LSM3D_TVD_RK3_STAGE3",
    "lsm3dcomputereinitializationeqnrhs_": "LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS",
    "lsm3dcomputeorthogonalizationeqnrhs_": "LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS",
    "lsm3dcomputedistanceforsubcellfix_": "LSM3D_COMPUTE_DISTANCE_FOR_SUBCELL_FIX",
    "lsm3dcomputereinitializationeqnrhssubcellfixorder1_":
        "LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_SUBCELL_FIX_ORDER1",
    "lsm3dmaxnormdiff_": "LSM3D_MAX_NORM_DIFF",
    "lsm3dmaxnormdiffcontrolvolume_": "LSM3D_MAX_NORM_DIFF_CONTROL_VOLUME",
    "lsm3dcomputestableadvectiondt_": "LSM3D_COMPUTE_STABLE_ADVECTION_DT",
    "lsm3dcomputestableadvectiondtcontrolvolume_":
        "LSM3D_COMPUTE_STABLE_ADVECTION_DT_CONTROL_VOLUME",
    "lsm3dcomputestablenormalveldt_": "LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT",
    "lsm3dcomputestableconstnormalveldt_": "LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT",
    "lsm3dvolumeintegralphilessthanzero_": "LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO",
    "lsm3dvolumeintegralphigreaterthanzero_": "LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO",
    "lsm3dvolumeregionphilessthanzero_": "LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO",
    "lsm3dvolumeregionphigreaterthanzero_": "LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO",
    "lsm3dlinearextrapolation_": "LSM3D_LINEAR_EXTRAPOLATION",
    "lsm3dsignedlinearextrapolation_": "LSM3D_SIGNED_LINEAR_EXTRAPOLATION",
    "lsm3dcopyextrapolation_": "LSM3D_COPY_EXTRAPOLATION",
    "lsm3dhomogeneousneumanneno1_": "LSM3D_HOMOGENEOUS_NEUMANN_ENO1",
    "lsm3dhomogeneousneumanneno2_": "LSM3D_HOMOGENEOUS_NEUMANN_ENO2",
    "lsm3dhomogeneousneumanneno3_": "LSM3D_HOMOGENEOUS_NEUMANN_ENO3",
    "lsm3dhomogeneousneumannweno5_": "LSM3D_HOMOGENEOUS_NEUMANN_WENO5",
    "lsm3ddeterminenarrowband_": "LSM3D_DETERMINE_NARROW_BAND",
    "lsm3dmarknarrowbandboundarylayer_": "LSM3D_MARK_NARROW_BAND_BOUNDARY_LAYER",
    "lsm3ddeterminenarrowbandfrommask_": "LSM3D_DETERMINE_NARROW_BAND_FROM_MASK",
    "lsm3dmultiplycutofflserhslocal_": "LSM3D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL",
    "lsm3dcheckouternarrowbandlayer_": "LSM3D_CHECK_OUTER_NARROW_BAND_LAYER",
    "lsm3dimposemasklocal_": "LSM3D_IMPOSE_MASK_LOCAL",
    "lsm3dcopydatalocal_": "LSM3D_COPY_DATA_LOCAL",
    "lsm3dhjeno1local_": "LSM3D_HJ_ENO1_LOCAL",
    "lsm3dhjeno2local_": "LSM3D_HJ_ENO2_LOCAL",
    "lsm3dcentralgradorder2local_": "LSM3D_CENTRAL_GRAD_ORDER2_LOCAL",
    "lsm3dcentralgradorder4local_": "LSM3D_CENTRAL_GRAD_ORDER4_LOCAL",
    "lsm3dlaplacianorder2local_": "LSM3D_LAPLACIAN_ORDER2_LOCAL",
    "lsm3dcomputeavegradphilocal_": "LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL",
    "lsm3dgradientmagnitudelocal_": "LSM3D_GRADIENT_MAGNITUDE_LOCAL",
    "lsm3drk1steplocal_": "LSM3D_RK1_STEP_LOCAL",
    "lsm3dtvdrk2stage1local_": "LSM3D_TVD_RK2_STAGE1_LOCAL",
    "lsm3dtvdrk2stage2local_": "LSM3D_TVD_RK2_STAGE2_LOCAL",
    "lsm3dtvdrk3stage1local_": "LSM3D_TVD_RK3_STAGE1_LOCAL",
    "lsm3dtvdrk3stage2local_": "LSM3D_TVD_RK3_STAGE2_LOCAL",
    "lsm3dtvdrk3stage3local_": "LSM3D_TVD_RK3_STAGE3_LOCAL",
    "lsm3dzerooutlevelseteqnrhslocal_": "LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL",
    "lsm3daddadvectiontermtolserhslocal_": "LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS_LOCAL",
    "lsm3daddnormalveltermtolserhslocal_": "LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL",
    "lsm3daddconstnormalveltermtolserhslocal_": "LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL",
    "lsm3daddconstcurvtermtolserhslocal_": "LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS_LOCAL",
    "lsm3daddconstprecomputedcurvtermtolserhslocal_": "LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS_LOCAL",
    "lsm3daddexternalandnormalveltermtolserhslocal_": "LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL",
    "lsm3dcomputereinitializationeqnrhslocal_": "LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_LOCAL",
    "lsm3dcomputeorthogonalizationeqnrhslocal_": "LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS_LOCAL",
    "lsm3dmaxnormdifflocal_": "LSM3D_MAX_NORM_DIFF_LOCAL",
    "lsm3dcomputestableadvectiondtlocal_": "LSM3D_COMPUTE_STABLE_ADVECTION_DT_LOCAL",
    "lsm3dcomputestablenormalveldtlocal_": "LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL",
    "lsm3dcomputestablenormalveldtcontrolvolumelocal_":
        "LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_CONTROL_VOLUME_LOCAL",
    "lsm3dcomputestableconstnormalveldtlocal_": "LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL",
}

# the narrow-band routines (index lists and marks), a subset of SIGNATURES,
# and the box routines (the rest)
LOCAL = {n for n, k in SIGNATURES.items() if "X" in k or "U" in k}
BOX_ROUTINES = sorted(set(SIGNATURES) - LOCAL)

assert set(MACRO_OF) == set(SIGNATURES)
