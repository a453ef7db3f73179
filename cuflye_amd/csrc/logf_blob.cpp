// logf_blob.cpp - embeds the build-time glibc-logf agreement table
// (gen_logf_table.cpp) into libflyehip.so as read-only data.
#ifndef FH_LOGF_TABLE_PATH
#error "FH_LOGF_TABLE_PATH must name the generated table"
#endif
#define FH_STR2(x) #x
#define FH_STR(x) FH_STR2(x)
asm(".section .rodata\n"
	".global fh_logf_table\n"
	".balign 16\n"
	"fh_logf_table:\n"
	".incbin \"" FH_STR(FH_LOGF_TABLE_PATH) "\"\n"
	".global fh_logf_table_end\n"
	"fh_logf_table_end:\n"
	".byte 0\n"
	".previous\n");
