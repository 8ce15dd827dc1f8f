// Fiber switch of the host emulation (x86-64 System V): saves the callee-saved registers on
// the current stack, stores the stack pointer to *from_sp and resumes the fiber whose stack
// pointer is to_sp.
asm(R"(
.text
.globl shim_fiber_switch
.type shim_fiber_switch,@function
shim_fiber_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
)");
