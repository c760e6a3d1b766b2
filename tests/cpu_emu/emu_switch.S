/* Fiber context switch of the CPU emulator (x86-64 System V).
 * emu_ctx_switch(void** save_sp, void* load_sp): saves the callee-saved registers on the current
 * stack, stores the stack pointer to *save_sp, switches to load_sp and restores from there.
 * A new fiber's stack holds six zero registers and the address of emu_fiber_entry. */
        .text
        .globl  emu_ctx_switch
        .type   emu_ctx_switch, @function
emu_ctx_switch:
        pushq   %rbp
        pushq   %rbx
        pushq   %r12
        pushq   %r13
        pushq   %r14
        pushq   %r15
        movq    %rsp, (%rdi)
        movq    %rsi, %rsp
        popq    %r15
        popq    %r14
        popq    %r13
        popq    %r12
        popq    %rbx
        popq    %rbp
        ret
        .size   emu_ctx_switch, .-emu_ctx_switch

        .globl  emu_fiber_entry
        .type   emu_fiber_entry, @function
emu_fiber_entry:
        /* the stack is 16-byte aligned here, as at a call site */
        call    emu_fiber_main
        ud2
        .size   emu_fiber_entry, .-emu_fiber_entry

        .section .note.GNU-stack,"",@progbits
