/* emission restatement: filled in below */
