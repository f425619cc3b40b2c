!! node_evolver_xml.F90 -- Fortran host program for the node evolver's method
!! set: reads the evolver's parameters from a Galacticus parameter file and
!! prints them; given a batch file, evolves the batch on the GPU (with an
!! optional wall-time limit turned into a System_Clock maximum), applies a
!! list of promote/merge events and writes the result.
!!
!! Usage: node_evolver_xml params.xml [batch.bin out.bin]
!! Batch (stream, native endianness): int32 n, nT, nZ; float64 lnTMin, dLnT,
!! lnZMin, dLnZ, wallSeconds (<= 0: no limit); params(48); table(nT*nZ);
!! y(n*30); nodePar(n*7); tStart(n); tEnd(n); timeStep(n); int32 nEvents;
!! nEvents x int32 (kind 1 = promote, 3 = merge; node; target, 1-based);
!! float64 ratioMajor.
!! Output: y(n*30), timeStep(n), int32 status(n), int32 return code.
program node_evolver_xml
  use :: Galacticus_AMD_ODE, only : nodeEvolverBatched, galVariableCount, galParameterCount, galNodeParameterCount
  implicit none
  type            (nodeEvolverBatched)          :: evolver
  integer                                       :: n, countT, countZ, countEvents, status, unitIn, unitOut, i
  integer         , allocatable, dimension(:,:) :: events
  double precision                              :: lnTMinimum, lnTStep, lnZMinimum, lnZStep, wallSeconds, ratioMajor
  double precision, allocatable, dimension(:  ) :: parameters, table, timeStart, timeEnd, timeStep
  double precision, allocatable, dimension(:,:) :: y, nodeParameters, scale
  integer         , allocatable, dimension(:  ) :: statusNode
  integer         (kind=8)                      :: clockCount, clockRate, clockMaximum
  character       (len=1024)                    :: fileParameters, fileIn, fileOut

  call get_command_argument(1,fileParameters)
  if (evolver%readParameters(trim(fileParameters)) /= 0) stop 2
  write (*,'(a,es24.16)') 'odeToleranceAbsolute ', evolver%odeToleranceAbsolute
  write (*,'(a,es24.16)') 'odeToleranceRelative ', evolver%odeToleranceRelative
  write (*,'(a,l1)'     ) 'reuseODEStepSize '    , evolver%reuseODEStepSize
  write (*,'(a,l1)'     ) 'enforceNonNegativity ', evolver%enforceNonNegativity
  if (command_argument_count() < 3) stop
  call get_command_argument(2,fileIn )
  call get_command_argument(3,fileOut)
  open(newunit=unitIn,file=trim(fileIn),access='stream',form='unformatted',status='old',action='read')
  read(unitIn) n, countT, countZ
  read(unitIn) lnTMinimum, lnTStep, lnZMinimum, lnZStep, wallSeconds
  allocate(parameters(galParameterCount),table(countT*countZ))
  allocate(y(n,galVariableCount),nodeParameters(n,galNodeParameterCount),scale(n,galVariableCount))
  allocate(timeStart(n),timeEnd(n),timeStep(n),statusNode(n))
  read(unitIn) parameters, table, y, nodeParameters, timeStart, timeEnd, timeStep
  read(unitIn) countEvents
  allocate(events(3,countEvents))
  read(unitIn) events, ratioMajor
  close(unitIn)
  status=evolver%initialize(n,parameters)
  if (status == 0) status=evolver%setCoolingTable(table,countT,lnTMinimum,lnTStep,countZ,lnZMinimum,lnZStep)
  scale     =0.0d0
  statusNode=0
  if (status == 0) then
     if (wallSeconds > 0.0d0) then
        call System_Clock(clockCount,clockRate)
        clockMaximum=clockCount+int(wallSeconds*dble(clockRate),kind=8)
        status=evolver%evolve(y,nodeParameters,timeStart,timeEnd,timeStep,scale,statusNode,clockMaximum)
     else
        status=evolver%evolve(y,nodeParameters,timeStart,timeEnd,timeStep,scale,statusNode)
     end if
  end if
  do i=1,countEvents
     if (status /= 0) exit
     select case (events(1,i))
     case (1)
        status=evolver%promote(y,timeStep,events(2,i),events(3,i))
     case (3)
        status=evolver%merge  (y,timeStep,events(2,i),events(3,i),ratioMajor)
     case default
        status=-1
     end select
  end do
  open(newunit=unitOut,file=trim(fileOut),access='stream',form='unformatted',status='replace',action='write')
  write(unitOut) y, timeStep, statusNode, status
  close(unitOut)
  if (evolver%destroy() /= 0 .or. status /= 0) stop 1
end program node_evolver_xml
