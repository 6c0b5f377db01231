/*
 * Status objects (status.h:9-17) are created by TAPP_AMD_server_submit and
 * released with TAPP_destroy_status (serve.c).  TAPP_execute_product is
 * synchronous and writes no status, like the reference.
 */
