/*
 * hfv_error.c -- ErrorStatus construction and tracebacks (the convention of
 * the reference's error.h:37-132).  A failing call returns an ErrorStatus whose
 * heap-allocated traceback string is owned by the caller and released with
 * free_traceback / print_and_free_traceback; each layer that forwards an
 * error prepends one frame with WRAP_TRACEBACK.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hydrofv/hydro.h"

ErrorStatus make_success_error_status(void)
{
    ErrorStatus s;
    s.return_code = SUCCESS;
    s.traceback = NULL;
    s.traceback_code_ = TRACEBACK_NOT_INITIALIZED;
    return s;
}

static const char *error_name(int code)
{
    switch (code) {
    case FAILURE: return "Failure";
    case VALUE_ERROR: return "ValueError";
    case POINTER_ERROR: return "PointerError";
    case MEMORY_ERROR: return "MemoryError";
    case OS_ERROR: return "OSError";
    case NOT_IMPLEMENTED_ERROR: return "NotImplementedError";
    default: return "UnknownError";
    }
}

void raise_warning(const char *warning_msg, const char *warning_file, const int warning_line,
                   const char *warning_func)
{
    fprintf(stderr, "Warning: In \"%s\", line %d in %s:\n    %s\n", warning_file, warning_line, warning_func,
            warning_msg);
}

/* snprintf into a fresh heap buffer; sets the traceback code accordingly */
static char *format_frame(int *code, const char *fmt, const char *a, int line, const char *b, const char *c,
                          const char *d)
{
    const int need = snprintf(NULL, 0, fmt, a, line, b, c, d);
    if (need < 0) {
        *code = TRACEBACK_SNPRINTF_FAILED;
        return NULL;
    }
    char *buf = malloc((size_t)need + 1);
    if (!buf) {
        *code = TRACEBACK_MALLOC_FAILED;
        return NULL;
    }
    snprintf(buf, (size_t)need + 1, fmt, a, line, b, c, d);
    *code = TRACEBACK_SUCCESS;
    return buf;
}

ErrorStatus raise_error(const int error_code, const char *error_msg, const char *error_file, const int error_line,
                        const char *error_func)
{
    ErrorStatus s;
    const char *name = error_name(error_code);
    s.return_code = (strcmp(name, "UnknownError") == 0) ? UNKNOWN_ERROR : error_code;
    s.traceback = format_frame(&s.traceback_code_, "    File \"%s\", line %d in %s\n%s: %s\n", error_file,
                               error_line, error_func, name, error_msg);
    return s;
}

ErrorStatus traceback(ErrorStatus error_status, const char *function_call_source_code, const char *error_file,
                      const int error_line, const char *error_func)
{
    if (error_status.return_code == SUCCESS || error_status.traceback_code_ != TRACEBACK_SUCCESS) {
        return error_status;
    }
    int code;
    char *frame = format_frame(&code, "    File \"%s\", line %d in %s\n        %s\n%s", error_file, error_line,
                               error_func, function_call_source_code, error_status.traceback);
    free(error_status.traceback);
    error_status.traceback = frame;
    error_status.traceback_code_ = code;
    return error_status;
}

void free_traceback(ErrorStatus *error_status)
{
    if (error_status->traceback) {
        free(error_status->traceback);
        error_status->traceback = NULL;
    }
}

void print_and_free_traceback(ErrorStatus *error_status)
{
    fputs("\nTraceback (most recent call last):\n", stderr);
    switch (error_status->traceback_code_) {
    case TRACEBACK_SUCCESS:
    case TRACEBACK_TRUNCATED:
        if (error_status->traceback) fputs(error_status->traceback, stderr);
        break;
    case TRACEBACK_MALLOC_FAILED:
        fputs("    Something went wrong. Failed to allocate memory for traceback.\n", stderr);
        break;
    case TRACEBACK_SNPRINTF_FAILED:
        fputs("    Something went wrong. Failed to write to traceback.\n", stderr);
        break;
    default:
        fputs("    Something went wrong. Traceback not initialized.\n", stderr);
        break;
    }
    free_traceback(error_status);
}
