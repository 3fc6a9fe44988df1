! emergency_stop with the reference's behaviour (src/error.f90:24-46):
! print the message, write it to ERROR, delete RUNNING, stop.
module error
  implicit none
  private
  public :: emergency_stop
contains
  subroutine emergency_stop(msg)
    character(len=*), intent(in) :: msg
    integer :: u
    logical :: ex
    print *, msg
    open (newunit=u, file='ERROR')
    write (u, *) msg
    close (u)
    inquire (file='RUNNING', exist=ex)
    if (ex) then
      open (newunit=u, file='RUNNING')
      close (u, status='delete')
    end if
    stop 1
  end subroutine emergency_stop
end module error
