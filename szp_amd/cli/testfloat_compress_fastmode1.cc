// Compresses a raw float/double file with szp_compress() (libszp_amd.so) and
// writes <file>.szp.  Arguments and printed metrics follow the reference's
// example/testfloat_compress_fastmode1.cc:
//   testfloat_compress_fastmode1 srcFilePath blockSize errBound {float|double}
//                                {random|nonrandom} [repetitions] [warmup_runs] [output_csv]
// The error bound is absolute (ABS).  T, which fixes the stream layout, is
// the library's default (SZP_AMD_NUM_THREADS, else OMP_NUM_THREADS, else the
// CPUs this process may use), as the reference takes its OpenMP team size.
#include "cli_common.h"

static void usage()
{
    printf("Usage: testfloat_compress_fastmode1 [srcFilePath] [block size] [err bound] [data_type] "
           "[access_type] [repetitions] [warmup_runs] [output_csv]\n");
    printf("Example: testfloat_compress_fastmode1 field.f32 128 1E-3 float nonrandom 10 2 results.csv\n");
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        usage();
        return 0;
    }
    const char *src = argv[1];
    const int blockSize = atoi(argv[2]);
    const double errBound = atof(argv[3]);
    const int dtype = cli::parse_data_type(argv[4]);
    const int access = cli::parse_access(argv[5]);
    const int reps = argc >= 7 ? atoi(argv[6]) : 1;
    const int warmup = argc >= 8 ? atoi(argv[7]) : 0;
    const char *csv = argc >= 9 ? argv[8] : nullptr;
    if (dtype < 0) {
        printf("Error: data type must be 'float' or 'double'!\n");
        return 0;
    }
    if (access < 0) {
        printf("Error: access type must be 'random' or 'nonrandom'!\n");
        return 0;
    }
    std::vector<char> out_path(strlen(src) + 8);
    snprintf(out_path.data(), out_path.size(), "%s.szp", src);

    printf("SZp Compression Benchmark Configuration\n");
    printf("======================================\n");
    printf("Input file: %s\nBlock size: %d\nError bound: %g\nData type: %s\nAccess type: %s\n",
           src, blockSize, errBound, argv[4], argv[5]);
    printf("Repetitions: %d\nWarmup runs: %d\n", reps, warmup);
    if (csv) printf("CSV output: %s\n", csv);
    printf("======================================\n\n");

    int status = SZ_SCES;
    size_t nbEle = 0;
    printf("Reading input data...\n");
    void *data = dtype == SZ_FLOAT ? (void *)szp_readFloatData((char *)src, &nbEle, &status)
                                   : (void *)szp_readDoubleData((char *)src, &nbEle, &status);
    if (status != SZ_SCES) {
        printf("Error: data file %s cannot be read!\n", src);
        return 0;
    }
    printf("Data elements: %zu\n", nbEle);
    const double raw = (double)nbEle * (dtype == SZ_FLOAT ? sizeof(float) : sizeof(double));

    if (warmup > 0) printf("\nPerforming %d warmup runs...\n", warmup);
    for (int i = 0; i < warmup; ++i) {
        size_t outSize = 0;
        const double t0 = cli::now_seconds();
        unsigned char *bytes = szp_compress(access, dtype, data, &outSize, ABS, errBound, 0,
                                            nbEle, blockSize);
        const double t = cli::now_seconds() - t0;
        printf("  Warmup run %d of %d: time=%f, size=%zu, CR=%f\n", i + 1, warmup, t, outSize,
               outSize ? raw / outSize : 0.0);
        free(bytes);
    }

    std::vector<double> times, ratios;
    std::vector<size_t> sizes;
    printf("\nStarting %d benchmark runs...\n", reps);
    for (int i = 0; i < reps; ++i) {
        printf("\nRun %d of %d:\n", i + 1, reps);
        size_t outSize = 0;
        const double t0 = cli::now_seconds();
        unsigned char *bytes = szp_compress(access, dtype, data, &outSize, ABS, errBound, 0,
                                            nbEle, blockSize);
        const double t = cli::now_seconds() - t0;
        if (!bytes) {
            printf("Error: compression failed\n");
            free(data);
            return 1;
        }
        const double cr = raw / outSize;
        printf("  timecost=%f, compression size=%zu, CR=%f\n", t, outSize, cr);
        times.push_back(t);
        ratios.push_back(cr);
        sizes.push_back(outSize);
        if (i == reps - 1) {
            printf("  Writing compressed data to %s\n", out_path.data());
            szp_writeByteData(bytes, outSize, out_path.data(), &status);
            if (status != SZ_SCES) printf("Error: data file %s cannot be written!\n", out_path.data());
        }
        free(bytes);
    }
    if (times.size() > 1) {
        printf("\n======= Performance Statistics (excluding warmup runs) =======\n");
        cli::print_stat("Time (seconds)", times);
        cli::print_stat("Compression Ratio", ratios);
        printf("===========================================================\n");
    }
    if (csv) {
        FILE *f = fopen(csv, "w");
        if (!f) {
            printf("Error: Cannot open file %s for writing\n", csv);
        } else {
            fprintf(f, "Run,TimeSeconds,CompressionSize,CompressionRatio\n");
            for (size_t i = 0; i < times.size(); ++i)
                fprintf(f, "%zu,%f,%zu,%f\n", i + 1, times[i], sizes[i], ratios[i]);
            fclose(f);
            printf("Results written to %s\n", csv);
        }
    }
    free(data);
    return 0;
}
