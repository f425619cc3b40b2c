!! galacticus_amd_trees.F90 -- Fortran host binding of the tree path.
!!
!! ISO_C_BINDING interfaces to include/gal_tree.h (forest builder, the
!! reference's tree walk recorded as a schedule, the wave executor on the GPU)
!! and to the four HIP runtime calls a Fortran host needs to keep the forest's
!! state in device memory.  The derived types mirror the C structs member for
!! member.  In the reference the tree evolver (merger_trees/evolver/standard.F90)
!! walks the forest and calls the node evolver per node; here the host builds
!! the walk once and hands the whole forest to the GPU.
module Galacticus_AMD_Trees
  use, intrinsic :: ISO_C_Binding, only : c_int, c_int32_t, c_int64_t, c_double, c_ptr, c_size_t, c_null_ptr
  implicit none
  private
  public :: galCosmology, galPCHParameters, galWalkParameters
  public :: gal_forest_build_pch, gal_forest_size, gal_forest_arrays, gal_forest_destroy, gal_forest_initial_state
  public :: gal_schedule_build, gal_schedule_size, gal_schedule_destroy, gal_tree_evolve_d
  public :: hipMalloc, hipFree, hipMemcpy, hipDeviceSynchronize, hipMemcpyHostToDevice, hipMemcpyDeviceToHost

  integer(c_int), parameter :: hipMemcpyHostToDevice=1, hipMemcpyDeviceToHost=2

  !! gal_cosmology
  type, bind(c) :: galCosmology
     real(c_double) :: omegaMatter, omegaDarkEnergy, omegaBaryon, hubbleConstant, sigma8, indexSpectral, temperatureCMB
  end type galCosmology

  !! gal_pch_params
  type, bind(c) :: galPCHParameters
     real(c_double) :: massResolution, mergeProbability, accretionLimit, g0, gamma1, gamma2, epsilon1, timeEarliest
  end type galPCHParameters

  !! gal_walk_params
  type, bind(c) :: galWalkParameters
     real(c_double) :: timeEnd, timestepRelative, timestepAbsolute, timestepHostRelative, timestepHostAbsolute, &
          &            satelliteOffsetAbsolute, satelliteOffsetRelative, mergingMultiplier, orbitRadiusRatio,   &
          &            orbitCircularity, spinParameter
  end type galWalkParameters

  interface
     function gal_forest_build_pch(cosmology,pch,countTrees,massRoot,timeRoot,seed,forest) &
          & bind(c,name='gal_forest_build_pch')
       import c_int, c_int64_t, c_double, c_ptr, galCosmology, galPCHParameters
       type   (galCosmology    )               :: cosmology
       type   (galPCHParameters)               :: pch
       integer(c_int           ), value        :: countTrees
       real   (c_double        ), dimension(*) :: massRoot
       real   (c_double        ), value        :: timeRoot
       integer(c_int64_t       ), value        :: seed
       type   (c_ptr           )               :: forest
       integer(c_int           )               :: gal_forest_build_pch
     end function gal_forest_build_pch
     function gal_forest_size(forest,countNodes,countTrees) bind(c,name='gal_forest_size')
       import c_int, c_int64_t, c_ptr
       type   (c_ptr    ), value :: forest
       integer(c_int64_t)        :: countNodes
       integer(c_int    )        :: countTrees
       integer(c_int    )        :: gal_forest_size
     end function gal_forest_size
     function gal_forest_arrays(forest,treeOffset,parent,firstChild,sibling,mass,time) bind(c,name='gal_forest_arrays')
       import c_int, c_int32_t, c_int64_t, c_double, c_ptr
       type   (c_ptr    ), value        :: forest
       integer(c_int64_t), dimension(*) :: treeOffset
       integer(c_int32_t), dimension(*) :: parent, firstChild, sibling
       real   (c_double ), dimension(*) :: mass, time
       integer(c_int    )               :: gal_forest_arrays
     end function gal_forest_arrays
     function gal_forest_destroy(forest) bind(c,name='gal_forest_destroy')
       import c_int, c_ptr
       type   (c_ptr), value :: forest
       integer(c_int)        :: gal_forest_destroy
     end function gal_forest_destroy
     function gal_forest_initial_state(forest,cosmology,walk,metalFraction,blackHoleSeed,ld,y) &
          & bind(c,name='gal_forest_initial_state')
       import c_int, c_double, c_ptr, galCosmology, galWalkParameters
       type   (c_ptr            ), value        :: forest
       type   (galCosmology     )               :: cosmology
       type   (galWalkParameters)               :: walk
       real   (c_double         ), value        :: metalFraction, blackHoleSeed
       integer(c_int            ), value        :: ld
       real   (c_double         ), dimension(*) :: y
       integer(c_int            )               :: gal_forest_initial_state
     end function gal_forest_initial_state
     function gal_schedule_build(forest,cosmology,walk,schedule) bind(c,name='gal_schedule_build')
       import c_int, c_ptr, galCosmology, galWalkParameters
       type   (c_ptr            ), value :: forest
       type   (galCosmology     )        :: cosmology
       type   (galWalkParameters)        :: walk
       type   (c_ptr            )        :: schedule
       integer(c_int            )        :: gal_schedule_build
     end function gal_schedule_build
     function gal_schedule_size(schedule,countOperations,countWaves) bind(c,name='gal_schedule_size')
       import c_int, c_int64_t, c_ptr
       type   (c_ptr    ), value :: schedule
       integer(c_int64_t)        :: countOperations
       integer(c_int    )        :: countWaves
       integer(c_int    )        :: gal_schedule_size
     end function gal_schedule_size
     function gal_schedule_destroy(schedule) bind(c,name='gal_schedule_destroy')
       import c_int, c_ptr
       type   (c_ptr), value :: schedule
       integer(c_int)        :: gal_schedule_destroy
     end function gal_schedule_destroy
     function gal_tree_evolve_d(evolver,schedule,ld,y,scale,status,majorRatio,stream) bind(c,name='gal_tree_evolve_d')
       import c_int, c_double, c_ptr
       type   (c_ptr   ), value :: evolver, schedule, y, scale, status, stream
       integer(c_int   ), value :: ld
       real   (c_double), value :: majorRatio
       integer(c_int   )        :: gal_tree_evolve_d
     end function gal_tree_evolve_d
     ! HIP runtime (device memory for the forest state)
     function hipMalloc(pointer,size) bind(c,name='hipMalloc')
       import c_int, c_ptr, c_size_t
       type   (c_ptr   )        :: pointer
       integer(c_size_t), value :: size
       integer(c_int   )        :: hipMalloc
     end function hipMalloc
     function hipFree(pointer) bind(c,name='hipFree')
       import c_int, c_ptr
       type   (c_ptr), value :: pointer
       integer(c_int)        :: hipFree
     end function hipFree
     function hipMemcpy(destination,source,size,kind) bind(c,name='hipMemcpy')
       import c_int, c_ptr, c_size_t
       type   (c_ptr   ), value :: destination, source
       integer(c_size_t), value :: size
       integer(c_int   ), value :: kind
       integer(c_int   )        :: hipMemcpy
     end function hipMemcpy
     function hipDeviceSynchronize() bind(c,name='hipDeviceSynchronize')
       import c_int
       integer(c_int) :: hipDeviceSynchronize
     end function hipDeviceSynchronize
  end interface

end module Galacticus_AMD_Trees
